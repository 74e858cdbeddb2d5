// realize_gap.hip -- on-GPU Poisson realizations for the loud/background split.
//
// Same outputs and selection rules as the injected-count kernel in
// realize.hip (reference: cyutils.pyx:862-897, 935-1178, 1266-1344, 1615-1767);
// this file draws the counts.
//
// Counts n_b ~ Poisson(lambda_b) are independent per (bin, frequency,
// realization).  Per frequency, bins are split by lambda:
//   * lambda == 0: n = 0, skipped;
//   * 0 < lambda < tau ("sparse"): the counts of the sparse bins of a tile are
//     the counts of one unit-rate Poisson process on the tile's cumulative-
//     lambda axis -- N ~ Poisson(sum lambda) points, each uniform on the axis,
//     assigned to the bin whose interval holds it.  This is exact (disjoint
//     intervals of a Poisson process have independent Poisson counts) and
//     costs O(sum lambda) per realization instead of O(#bins);
//   * lambda >= tau ("dense"): one draw per bin: inversion below 10, PTRS
//     (Hoermann 1993) at and above 10, Normal(lambda, sqrt(lambda)) above the
//     normal threshold, as the reference (cyutils.pyx:1328-1332).
//
// Two kernels:
//   k_gz_prep  -- per (tile of 512 bins, frequency): classify, compact into
//                 per-tile dense records (sampler constants precomputed once
//                 for all realizations) and sparse records (cumulative lambda);
//   k_gz_draw  -- per (chunk of tiles, frequency): every realization (one per
//                 thread, one Philox block per dense draw attempt); dense
//                 records are read with wave-uniform loads, sparse points are
//                 placed by binary search on the tile's cumulative axis.
// Per-(frequency, chunk, realization) partials (FP64 totals and the L best
// candidates) are merged by realize.hip's merge kernel.
#include "common.hpp"
#include "rng.hpp"
#include "../../include/holohip.h"

namespace hd {

constexpr int GZ_TILE = 512;          // bins per tile (prep workgroup size)
constexpr int GZ_TILES_PER_CHUNK = 16;
constexpr int GZ_THREADS = 256;       // draw workgroup: one realization per thread
constexpr uint64_t GZ_KEY_NONE = ~0ull;

// dense record: 8 doubles
struct DRec {
    double lam, h, c0, c1, c2, c3, c4;
    uint64_t word;   // bin << 32 | rank
};

struct GapArgs {
    const double* number;   // (B,F)
    const double* h2fdf;    // (B,F)
    const int32_t* rank;    // (B) order rank (SORTED) or null
    const int64_t* order;   // (B) (SORTED) or null
    const double* par;      // (npar,B,F) or null
    int64_t B;
    int F, R, L, flags, npar, ntiles, nchunks, mode;
    double thresh, tau;
    uint64_t seed;
    // tile records, region (f, tile) holds up to GZ_TILE entries
    DRec* drec;             // (F, ntiles, GZ_TILE)
    double* s_cum;          // (F, ntiles, GZ_TILE) inclusive cumulative lambda
    double* s_h;            // (F, ntiles, GZ_TILE)
    uint64_t* s_word;       // (F, ntiles, GZ_TILE)
    double* d_par;          // (npar, F, ntiles, GZ_TILE)
    double* sp_par;         // (npar, F, ntiles, GZ_TILE)
    int32_t* cnt;           // (F, ntiles, 2): dense, sparse counts
    // partials
    double* p_total;        // (F, nchunks, R)
    double* p_ptot;         // (npar, F, nchunks, R)
    uint64_t* p_key;        // (F, nchunks, R, L)
    int32_t* p_bin;         // (F, nchunks, R, L)
};

__device__ __forceinline__ uint64_t gz_hkey(double h) {
    uint64_t u = (uint64_t)__double_as_longlong(h);
    u = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return ~u;
}

// ---------------------------------------------------------------------------
// prep: classify and compact one tile of one frequency
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(GZ_TILE) k_gz_prep(GapArgs A) {
    __shared__ int s_wd[GZ_TILE / 64], s_ws[GZ_TILE / 64];
    __shared__ double s_wsum[GZ_TILE / 64];
    const int f = blockIdx.y, tile = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int F = A.F;
    const int64_t b = (int64_t)tile * GZ_TILE + tid;
    const bool skip_zero_h = (A.flags & HD_FLAG_SKIP_ZERO_H) != 0;
    double lam = 0.0, h = 0.0;
    bool in = b < A.B;
    if (in) {
        lam = A.number[b * F + f];
        h = A.h2fdf[b * F + f];
        if (!(lam > 0.0)) in = false;
        if (A.mode == HD_SEL_SORTED && skip_zero_h && h == 0.0) in = false;
        // argmax: h <= 0 adds n*h = 0 and can never be selected
        if (A.mode == HD_SEL_ARGMAX && !(h > 0.0)) in = false;
    }
    const bool dense = in && (lam >= A.tau);
    const bool sparse = in && !dense;
    const uint64_t md = __ballot(dense), ms = __ballot(sparse);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    double x = sparse ? lam : 0.0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    if (lane == 0) {
        s_wd[wave] = __popcll(md);
        s_ws[wave] = __popcll(ms);
    }
    __syncthreads();
    int od = 0, os = 0;
    double osum = 0.0;
    for (int w = 0; w < wave; w++) {
        od += s_wd[w];
        os += s_ws[w];
        osum += s_wsum[w];
    }
    const int64_t reg = ((int64_t)f * A.ntiles + tile) * GZ_TILE;
    const uint32_t rk = (A.mode == HD_SEL_SORTED && in) ? (uint32_t)A.rank[b] : 0u;
    const uint64_t word = ((uint64_t)(uint32_t)b << 32) | rk;
    if (dense) {
        const int i = od + __popcll(md & lt);
        DRec r;
        r.lam = lam;
        r.h = h;
        r.c0 = r.c1 = r.c2 = r.c3 = r.c4 = 0.0;
        if (lam > A.thresh) {
            r.c0 = sqrt(lam);
        } else if (lam < 10.0) {
            r.c0 = -expm1(-lam);
            r.c1 = exp(-lam);
        } else {
            const double slam = sqrt(lam);
            const double bb = 0.931 + 2.53 * slam;
            const double aa = -0.059 + 0.02483 * bb;
            const double invalpha = 1.1239 + 1.1328 / (bb - 3.4);
            r.c0 = bb;
            r.c1 = aa;
            r.c2 = log(invalpha);
            r.c3 = 0.9277 - 3.6224 / (bb - 2);
            r.c4 = log(lam);
        }
        r.word = word;
        A.drec[reg + i] = r;
        for (int k = 0; k < A.npar; k++)
            A.d_par[(int64_t)k * F * A.ntiles * GZ_TILE + reg + i] = A.par[((int64_t)k * A.B + b) * F + f];
    }
    if (sparse) {
        const int i = os + __popcll(ms & lt);
        A.s_cum[reg + i] = osum + x;
        A.s_h[reg + i] = h;
        A.s_word[reg + i] = word;
        for (int k = 0; k < A.npar; k++)
            A.sp_par[(int64_t)k * F * A.ntiles * GZ_TILE + reg + i] = A.par[((int64_t)k * A.B + b) * F + f];
    }
    if (tid == 0) {
        int nd = 0, ns = 0;
        for (int w = 0; w < GZ_TILE / 64; w++) {
            nd += s_wd[w];
            ns += s_ws[w];
        }
        A.cnt[(((int64_t)f * A.ntiles) + tile) * 2] = nd;
        A.cnt[(((int64_t)f * A.ntiles) + tile) * 2 + 1] = ns;
    }
}

// ---------------------------------------------------------------------------
// draw
// ---------------------------------------------------------------------------
// Per-realization candidate list of the LM best instances; the worst of the
// first L slots is cached.  The sorted walk ranks instances by their position
// in the order (32-bit key; the bin is recovered from the order at the end);
// argmax ranks by (h descending, bin ascending).
template <int MODE, int LM>
struct LoudList;

template <int LM>
struct LoudList<HD_SEL_SORTED, LM> {
    uint32_t k[LM];
    uint32_t wk;
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < LM; i++) k[i] = 0xffffffffu;
        wk = 0xffffffffu;
    }
    __device__ __forceinline__ bool admits(uint64_t word, double) const { return (uint32_t)word < wk; }
    __device__ __forceinline__ void insert(uint64_t word, double, int L) {
        uint32_t key = (uint32_t)word;
#pragma unroll
        for (int i = 0; i < LM; i++) {
            uint32_t t = k[i];
            bool sw = key < t;
            k[i] = sw ? key : t;
            key = sw ? t : key;
        }
#pragma unroll
        for (int i = 0; i < LM; i++)
            if (i == L - 1) wk = k[i];
    }
    __device__ __forceinline__ uint64_t key_at(int j) const {
        return k[j] == 0xffffffffu ? GZ_KEY_NONE : (uint64_t)k[j];
    }
    __device__ __forceinline__ int32_t bin_at(int j, const int64_t* order) const {
        return k[j] == 0xffffffffu ? 0x7fffffff : (int32_t)order[k[j]];
    }
};

template <int MODE, int LM>
struct LoudList {
    uint64_t k[LM];
    int32_t b[LM];
    uint64_t wk;
    int32_t wb;
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < LM; i++) {
            k[i] = GZ_KEY_NONE;
            b[i] = 0x7fffffff;
        }
        wk = GZ_KEY_NONE;
        wb = 0x7fffffff;
    }
    __device__ __forceinline__ bool admits(uint64_t word, double h) const {
        const uint64_t key = gz_hkey(h);
        const int32_t bin = (int32_t)(word >> 32);
        return (key < wk) || (key == wk && bin < wb);
    }
    __device__ __forceinline__ void insert(uint64_t word, double h, int L) {
        uint64_t key = gz_hkey(h);
        int32_t bin = (int32_t)(word >> 32);
#pragma unroll
        for (int i = 0; i < LM; i++) {
            bool sw = (key < k[i]) || (key == k[i] && bin < b[i]);
            uint64_t tk = k[i];
            int32_t tb = b[i];
            k[i] = sw ? key : tk;
            b[i] = sw ? bin : tb;
            key = sw ? tk : key;
            bin = sw ? tb : bin;
        }
#pragma unroll
        for (int i = 0; i < LM; i++)
            if (i == L - 1) {
                wk = k[i];
                wb = b[i];
            }
    }
    __device__ __forceinline__ uint64_t key_at(int j) const { return k[j]; }
    __device__ __forceinline__ int32_t bin_at(int j, const int64_t*) const { return b[j]; }
};

// inversion of the upper tail for 0 < lambda < 10 with a 53-bit uniform
__device__ __forceinline__ double gz_inversion(double u, double lam, double q1, double elam) {
    int k = 0;
    double S = q1, p = elam;
    while (u < S && k < 200) {
        k++;
        p = p * lam / k;
        S = S - p;
    }
    return (double)k;
}

// PTRS acceptance with a uniform pair (U in (-0.5,0.5), V); returns k or -1 (reject)
__device__ __forceinline__ double gz_ptrs_try(double U, double V, const DRec& r) {
    const double us = 0.5 - fabs(U);
    const double k = floor((2 * r.c1 / us + r.c0) * U + r.lam + 0.43);
    if ((us >= 0.07) && (V <= r.c3)) return k;
    if ((k < 0) || ((us < 0.013) && (V > us))) return -1.0;
    const bool acc = (log(V) + r.c2 - log(r.c1 / (us * us) + r.c0)) <= (-r.lam + k * r.c4 - lgamma(k + 1.0));
    return acc ? k : -1.0;
}

// Count of one dense bin for realization r: Philox counter (bin, f, r,
// tag 1 | attempt).  Inversion uses words (x,y); a PTRS attempt uses (x,y) for
// U and (z,w) for V; the Normal branch uses both pairs for Box-Muller.
__device__ __forceinline__ double gz_dense_draw(const DRec& rec, uint32_t f, uint32_t r, uint64_t seed, double thresh) {
    const uint32_t bin = (uint32_t)(rec.word >> 32);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t TAG = 1u << 16;
    U4 a = philox4x32_10(U4{bin, f, r, TAG}, k0, k1);
    if (rec.lam < 10.0) return gz_inversion(u53(a.x, a.y), rec.lam, rec.c0, rec.c1);
    if (rec.lam > thresh)
        return rec.lam + rec.c0 * (sqrt(-2.0 * log(u53(a.x, a.y))) * cos(2.0 * M_PI * u53(a.z, a.w)));
    double n = gz_ptrs_try(u53(a.x, a.y) - 0.5, u53(a.z, a.w), rec);
#pragma unroll 1
    for (uint32_t att = 1; n < 0 && att < 4000u; att++) {
        a = philox4x32_10(U4{bin, f, r, TAG | att}, k0, k1);
        n = gz_ptrs_try(u53(a.x, a.y) - 0.5, u53(a.z, a.w), rec);
    }
    return (n < 0) ? floor(rec.lam) : n;   // the bound is unreachable in practice
}

template <int MODE, int LM, int NP>
struct Acc {
    LoudList<MODE, LM> list;
    double tot;
    double pt[NP > 0 ? NP : 1];
    __device__ __forceinline__ void init() {
        list.init();
        tot = 0.0;
#pragma unroll
        for (int k = 0; k < (NP > 0 ? NP : 1); k++) pt[k] = 0.0;
    }
    // n instances of a bin with strain h (n as drawn; the sorted walk ignores n < 1)
    __device__ __forceinline__ void add(double n, double h, uint64_t word, const double* p, int64_t pstride, int L) {
        if (MODE == HD_SEL_SORTED ? !(n >= 1.0) : (n == 0.0)) return;
        tot += n * h;
#pragma unroll
        for (int k = 0; k < NP; k++) pt[k] += n * h * p[k * pstride];
        if (MODE != HD_SEL_NONE && n > 0 && list.admits(word, h)) {
            const int ninst = (int)fmin(ceil(n), (double)L);
            for (int i = 0; i < ninst && list.admits(word, h); i++) list.insert(word, h, L);
        }
    }
};

template <int MODE, int LM, int NP>
__global__ void __launch_bounds__(GZ_THREADS) k_gz_draw(GapArgs A) {
    const int f = blockIdx.y;
    const int chunk = blockIdx.x;
    const int tid = threadIdx.x;
    const int F = A.F, R = A.R, L = A.L;
    const int t0 = chunk * GZ_TILES_PER_CHUNK;
    const int t1 = min(t0 + GZ_TILES_PER_CHUNK, A.ntiles);
    const int64_t pstride = (int64_t)F * A.ntiles * GZ_TILE;

    for (int rbase = 0; rbase < R; rbase += GZ_THREADS) {
        const int r = rbase + tid;
        const bool active = r < R;
        Acc<MODE, LM, NP> acc;
        acc.init();
        for (int tile = t0; tile < t1; tile++) {
            const int64_t ft = (int64_t)f * A.ntiles + tile;
            const int64_t reg = ft * GZ_TILE;
            const int nd = A.cnt[ft * 2];
            const int ns = A.cnt[ft * 2 + 1];
            // dense bins: wave-uniform loop, record in scalar registers
            for (int c = 0; c < nd; c++) {
                const DRec rec = A.drec[reg + c];
                const double n = gz_dense_draw(rec, (uint32_t)f, (uint32_t)r, A.seed, A.thresh);
                if (active) acc.add(n, rec.h, rec.word, A.d_par + reg + c, pstride, L);
            }
            // sparse bins: Poisson process on the tile's cumulative-lambda axis
            if (ns > 0 && active) {
                const double* cum = A.s_cum + reg;
                const double Lam = cum[ns - 1];
                Stream st;
                st.init((uint32_t)tile, (uint32_t)f, (uint32_t)r, 2u, A.seed);
                double npts;
                if (Lam < 10.0) npts = poisson_inversion(st.next(), Lam, -expm1(-Lam), exp(-Lam));
                else npts = poisson_ptrs(st, Lam);
                const int np = (int)npts;
                for (int i = 0; i < np; i++) {
                    const double x = st.next() * Lam;
                    int lo = 0, hi = ns - 1;
                    while (lo < hi) {   // first c with cum[c] > x
                        const int mid = (lo + hi) >> 1;
                        if (cum[mid] > x) hi = mid;
                        else lo = mid + 1;
                    }
                    acc.add(1.0, A.s_h[reg + lo], A.s_word[reg + lo], A.sp_par + reg + lo, pstride, L);
                }
            }
        }
        if (active) {
            const int64_t pr = ((int64_t)f * A.nchunks + chunk) * R + r;
            A.p_total[pr] = acc.tot;
#pragma unroll
            for (int k = 0; k < NP; k++) A.p_ptot[(int64_t)k * F * A.nchunks * R + pr] = acc.pt[k];
            if (MODE != HD_SEL_NONE) {
#pragma unroll
                for (int j = 0; j < LM; j++) {
                    if (j >= L) continue;
                    A.p_key[pr * L + j] = acc.list.key_at(j);
                    A.p_bin[pr * L + j] = acc.list.bin_at(j, A.order);
                }
            }
        }
    }
}

int realize_merge_launch(const double* h2fdf, const double* par, int64_t B, int F, int R, int L, int mode, int npar,
                         int nchunks, const double* p_total, const double* p_ptot, const uint64_t* p_key,
                         const int32_t* p_bin, double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar,
                         hipStream_t st);
int rank_from_order_launch(const int64_t* order, int64_t B, int32_t* rank, hipStream_t st);

static size_t gz_align(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t gz_ws(int64_t B, int F, int R, int L, int npar) {
    const int64_t nt = cdiv(B, GZ_TILE);
    const int64_t nch = cdiv(nt, GZ_TILES_PER_CHUNK);
    const size_t slots = (size_t)F * nt * GZ_TILE;
    const size_t pr = (size_t)F * nch * R;
    size_t s = gz_align((size_t)B * 4);                    // rank
    s += gz_align(slots * sizeof(DRec));                   // dense records
    s += 3 * gz_align(slots * 8);                          // sparse cum, h, word
    s += 2 * gz_align(slots * 8 * (size_t)npar);           // params
    s += gz_align((size_t)F * nt * 2 * 4);                 // counts
    s += gz_align(pr * 8) + gz_align(pr * 8 * (size_t)npar) + gz_align(pr * 8 * (size_t)L) +
         gz_align(pr * 4 * (size_t)L);
    return s + 512;
}

template <int MODE, int LM>
static void gz_launch_np(const GapArgs& A, int npar, dim3 grid, hipStream_t st) {
    if (npar == 0) hipLaunchKernelGGL((k_gz_draw<MODE, LM, 0>), grid, dim3(GZ_THREADS), 0, st, A);
    else if (npar <= 3) hipLaunchKernelGGL((k_gz_draw<MODE, LM, 3>), grid, dim3(GZ_THREADS), 0, st, A);
    else hipLaunchKernelGGL((k_gz_draw<MODE, LM, 7>), grid, dim3(GZ_THREADS), 0, st, A);
}

}  // namespace hd

using namespace hd;

extern "C" size_t hd_realize_workspace_size(int64_t B, int F, int R, int L, int npar) {
    return gz_ws(B, F, R, L, npar);
}

extern "C" int hd_realize(const double* number, const double* h2fdf, const int64_t* order, int64_t B, int F, int R, int L,
                          int mode, int flags, double normal_threshold, uint64_t seed, const double* par, int npar,
                          double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar, void* workspace,
                          size_t ws_bytes, void* stream) {
    if (B < 1 || F < 1 || R < 1 || !number || !h2fdf || !hc2bg) return set_error(HD_ERR_ARG, "hd_realize: bad arguments");
    if (B > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_realize: B=%lld exceeds 2^31-1", (long long)B);
    if (F > 65535) return set_error(HD_ERR_SIZE, "hd_realize: F=%d > 65535", F);
    if (mode != HD_SEL_SORTED && mode != HD_SEL_ARGMAX && mode != HD_SEL_NONE)
        return set_error(HD_ERR_ARG, "hd_realize: unknown mode %d", mode);
    if (mode == HD_SEL_NONE) L = 0;
    else if (L < 1) return set_error(HD_ERR_ARG, "hd_realize: L must be >= 1 for loud-source modes");
    if (L > 16) return set_error(HD_ERR_SIZE, "hd_realize: L=%d > 16", L);
    if (npar < 0 || npar > 7 || (npar > 0 && !par)) return set_error(HD_ERR_ARG, "hd_realize: npar=%d invalid", npar);
    if (mode == HD_SEL_SORTED && !order) return set_error(HD_ERR_ARG, "hd_realize: sorted mode needs an order");
    if (mode != HD_SEL_NONE && (!hc2ss || !ssidx)) return set_error(HD_ERR_ARG, "hd_realize: hc2ss/ssidx required");
    const size_t need = gz_ws(B, F, R, L, npar);
    if (!workspace || ws_bytes < need)
        return set_error(HD_ERR_SIZE, "hd_realize: workspace too small (%zu < %zu)", ws_bytes, need);
    hipStream_t st = (hipStream_t)stream;
    GapArgs A;
    A.number = number;
    A.h2fdf = h2fdf;
    A.order = order;
    A.par = par;
    A.B = B;
    A.F = F;
    A.R = R;
    A.L = L;
    A.flags = flags;
    A.npar = npar;
    A.mode = mode;
    A.thresh = normal_threshold;
    A.tau = 1.0;
    A.seed = seed;
    A.ntiles = (int)cdiv(B, GZ_TILE);
    A.nchunks = (int)cdiv(A.ntiles, GZ_TILES_PER_CHUNK);
    const size_t slots = (size_t)F * A.ntiles * GZ_TILE;
    const size_t pr = (size_t)F * A.nchunks * R;
    char* p = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    int32_t* rank = (int32_t*)p;  p += gz_align((size_t)B * 4);
    A.drec = (DRec*)p;            p += gz_align(slots * sizeof(DRec));
    A.s_cum = (double*)p;         p += gz_align(slots * 8);
    A.s_h = (double*)p;           p += gz_align(slots * 8);
    A.s_word = (uint64_t*)p;      p += gz_align(slots * 8);
    A.d_par = (double*)p;         p += gz_align(slots * 8 * (size_t)npar);
    A.sp_par = (double*)p;        p += gz_align(slots * 8 * (size_t)npar);
    A.cnt = (int32_t*)p;          p += gz_align((size_t)F * A.ntiles * 2 * 4);
    A.p_total = (double*)p;       p += gz_align(pr * 8);
    A.p_ptot = (double*)p;        p += gz_align(pr * 8 * (size_t)npar);
    A.p_key = (uint64_t*)p;       p += gz_align(pr * 8 * (size_t)L);
    A.p_bin = (int32_t*)p;
    A.rank = (mode == HD_SEL_SORTED) ? rank : nullptr;
    if (mode == HD_SEL_SORTED) {
        int rc = rank_from_order_launch(order, B, rank, st);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_gz_prep, dim3(A.ntiles, F), dim3(GZ_TILE), 0, st, A);
    int rc = check_launch("k_gz_prep");
    if (rc) return rc;
    const dim3 grid(A.nchunks, F);
    if (mode == HD_SEL_SORTED) {
        if (L <= 1) gz_launch_np<HD_SEL_SORTED, 1>(A, npar, grid, st);
        else if (L <= 4) gz_launch_np<HD_SEL_SORTED, 4>(A, npar, grid, st);
        else gz_launch_np<HD_SEL_SORTED, 16>(A, npar, grid, st);
    } else if (mode == HD_SEL_ARGMAX) {
        if (L <= 1) gz_launch_np<HD_SEL_ARGMAX, 1>(A, npar, grid, st);
        else if (L <= 4) gz_launch_np<HD_SEL_ARGMAX, 4>(A, npar, grid, st);
        else gz_launch_np<HD_SEL_ARGMAX, 16>(A, npar, grid, st);
    } else {
        gz_launch_np<HD_SEL_NONE, 1>(A, npar, grid, st);
    }
    rc = check_launch("k_gz_draw");
    if (rc) return rc;
    return realize_merge_launch(h2fdf, par, B, F, R, L, mode, npar, A.nchunks, A.p_total, A.p_ptot, A.p_key, A.p_bin,
                                hc2ss, ssidx, hc2bg, (npar > 0) ? bgpar : nullptr, st);
}
