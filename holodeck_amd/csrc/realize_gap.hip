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
//   * lambda >= tau ("dense"): one draw per bin and realization.
//
// Every bin is drawn R times (once per realization), so instead of a
// rejection sampler per draw the workgroup builds the bin's CDF once, in LDS,
// and every realization inverts its uniform by a branch-free binary search.
// With b = 56 ln 2 the support considered is [klo, khi],
//   klo = lambda - sqrt(2 b lambda)              (Chernoff lower tail:
//         P(X <= lambda - x) <= exp(-x^2 / 2 lambda))
//   khi = lambda + b/3 + sqrt(b^2/9 + 2 b lambda) (Bernstein upper tail:
//         P(X >= lambda + x) <= exp(-x^2 / (2 (lambda + x/3))))
// so each dropped tail holds less than 2^-56, below the 2^-64 resolution of the
// uniforms.  The table holds only the centre of it, counts kt .. kc with
// kt + 1 >= lambda - 4 sqrt(lambda) and kc >= lambda + 4 sqrt(lambda) + 1 (about
// 8 sqrt(lambda) entries instead of 17.6 sqrt(lambda)); its first entry is the
// whole CDF up to kt (the lower tail is summed, not stored).  A uniform below the
// first entry or above the last (probability about 1e-4 per draw) is resolved by
// walking the pmf recurrence from the table's edge, in FP64, down to klo or up to
// khi.  Bins whose table does not fit the LDS, and lambda above the normal
// threshold, are drawn by k_gz_big afterwards (PTRS, Hoermann 1993;
// Normal(lambda, sqrt(lambda)) unfloored as cyutils.pyx:1328-1332), which
// continues the same partials.  The sparse tile total N uses the same CDF
// machinery (as a pseudo-record appended to the tile; its lambda is below
// 512 tau, so its table always fits).
//
// Random numbers: each (tile, frequency, realization) has one xoshiro128**
// stream seeded by Philox4x32-10 of (tile, f, r, tag 6) under the seed; the
// tile's records take its words in order (gz_table_draw32), so results do not
// depend on launch shape or batching.
//
// Four kernels:
//   k_gz_prep  -- per (tile of 512 bins, group of 8 frequencies): classify,
//                 compact into per-tile dense records (lambda, h, key, CDF
//                 window and its offset in the tile's concatenated tables) and
//                 sparse records (cumulative lambda); the tile's cost estimate;
//   k_gz_plan  -- one workgroup: the draw workgroups' dispatch order, (chunk,
//                 frequency) items by decreasing estimated cost;
//   k_gz_draw  -- per (chunk of 8 tiles, frequency) workgroup, one realization
//                 per thread: tables of consecutive records are built in LDS
//                 in batches that fit (a lane per short table, a 16-lane row
//                 per mid-size table, a wave per long one); then every
//                 realization draws from them;
//   k_gz_big   -- per (chunk, frequency), only where the chunk has records
//                 without a table: reloads the partials, draws those records,
//                 writes the partials back.
// Per-(frequency, chunk, realization) partials (FP64 totals and the L best
// candidates) are merged by realize.hip's two merge kernels (chunk groups, then all).
#include "common.hpp"
#include "rng.hpp"
#include "../../include/holohip.h"
#include <mutex>

namespace hd {

constexpr int GZ_TILE = 512;          // bins per tile (prep workgroup size)
constexpr int GZ_REC = GZ_TILE + 1;   // record slots per tile (+1 sparse pseudo-record)
constexpr int GZ_TILES_PER_CHUNK = 8;    // items of 8 tiles: short enough to pack the GPU to its end
constexpr int GZ_MAXBS = 512;         // draw workgroup: one realization per thread
constexpr int GZ_BATCH_REC = 128;     // at most this many records per LDS batch (their headers are staged)
constexpr int GZ_XCDS = 8;            // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr uint64_t GZ_KEY_NONE = ~0ull;
constexpr double GZ_TAILB = 38.816242111356935;   // 56 ln 2: each dropped tail < 2^-56
constexpr double GZ_CENTRE = 4.0;                 // half-width of a table in standard deviations
#ifndef HD_GZ_TAU
// lambda below which a bin joins its tile's sparse Poisson process: a sparse bin costs about
// lambda point placements per realization, a dense one a table and a draw; the draw launch
// is fastest for tau in 0.35 .. 0.7 (measured at 91x81x101, F=40, R=500)
#define HD_GZ_TAU 0.5
#endif

// dense record (and the sparse pseudo-record at index nd)
struct DRec {
    double lam, h;
    uint64_t word;    // bin << 32 | rank
    double p0;        // Poisson pmf at kb
    int32_t kb;       // first count of the support (klo): the table build starts here
    int32_t n;        // table length (0: no table, use the fallback sampler)
    int32_t off;      // offset of the table in the tile's concatenated tables
    int32_t kt;       // count of the table's first entry
};

// header of a record staged in LDS for the draw loop (off relative to the batch's table)
struct BRec {
    double h;
    uint64_t word;
    double lam;
    int32_t kt, n, off, kb;
    double p0;   // pmf at kb (the build's start)
};

struct GapArgs {
    const double* number;   // (B,F)
    const double* h2fdf;    // (B,F)
    const int32_t* rank;    // (B) order rank (SORTED) or null
    const int64_t* order;   // (B) (SORTED) or null
    const double* par;      // (npar,B,F) or null
    int64_t B;
    int F, R, L, flags, npar, ntiles, nchunks, mode, cap;
    double thresh, tau;
    uint64_t seed;
    DRec* drec;             // (F, ntiles, GZ_REC)
    int32_t* rec_end;       // (F, ntiles, GZ_REC) off + n, nondecreasing within a tile
    int32_t* cnt;           // (F, ntiles, 3): dense, sparse, no-table counts
    double* s_cum;          // (F, ntiles, GZ_TILE) inclusive cumulative lambda
    double* s_h;            // (F, ntiles, GZ_TILE)
    uint64_t* s_word;       // (F, ntiles, GZ_TILE)
    double* d_par;          // (npar, F, ntiles, GZ_TILE)
    double* sp_par;         // (npar, F, ntiles, GZ_TILE)
    // partials
    double* p_total;        // (F, nchunks, R)
    double* p_ptot;         // (npar, F, nchunks, R)
    uint64_t* p_key;        // (F, nchunks, L, R): realization fastest; 32-bit ranks for the sorted walk
    int32_t* p_bin;         // (F, nchunks, L, R), argmax only
    double* counts_out;     // (B, F, R) every drawn count (test hook only; null in production)
    int32_t* plan;          // (nchunks * F) draw workgroup -> (chunk, f) item, f * nchunks + chunk
    int32_t* plan_b;        // (nchunks * F) the item's cost bucket
    float* tile_cost;       // (F, ntiles) estimated draw time of the tile (prep)
};

__device__ __forceinline__ uint64_t gz_hkey(double h) {
    uint64_t u = (uint64_t)__double_as_longlong(h);
    u = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return ~u;
}

// Support [klo, khi] of Poisson(lam) that drops less than 2^-56 on each side (as doubles).
__host__ __device__ __forceinline__ void gz_support(double lam, double& lo, double& hi) {
    const double b = GZ_TAILB;
    lo = fmax(floor(lam - sqrt(2.0 * b * lam)), 0.0);
    hi = ceil(lam + b / 3.0 + sqrt(b * b / 9.0 + 2.0 * b * lam));
}

// Table of Poisson(lam): the build starts at kb = klo, the table holds the counts
// [kt, kt + n) (entry 0 = the CDF up to kt), kt >= kb and kt + n - 1 <= khi.  n = 0
// (no table: the fallback sampler) if it exceeds cap entries, or if khi would not
// fit an int (lam near 2^31: the support would then collapse to a single node by
// rounding, and the counts overflow).
__host__ __device__ __forceinline__ void gz_window(double lam, int cap, int& kb, int& kt, int& n) {
    double lo, hi;
    gz_support(lam, lo, hi);
    const double sd = sqrt(lam);
    const double clo = fmax(floor(lam - GZ_CENTRE * sd), lo);           // first count of the centre
    const double chi = fmin(ceil(lam + GZ_CENTRE * sd) + 1.0, hi);      // last count of the table
    const double t0 = fmax(clo - 1.0, lo);                              // entry 0: the CDF below the centre
    const double w = chi - t0 + 1.0;
    if (!(w <= (double)cap) || !(hi < 2147483647.0)) {
        kb = kt = n = 0;
    } else {
        kb = (int)lo;
        kt = (int)t0;
        n = (int)w;
    }
}

// Stirling-formula error ln(k!) - (k + 1/2) ln k + k - ln sqrt(2 pi): k = 1..15 (from
// the definition at 40 digits), series above (error < 2e-16 at k = 16).
__constant__ double gz_stirlerr[16] = {
    0.0, 0.081061466795327258220, 0.041340695955409294094, 0.027677925684998339149, 0.020790672103765093112,
    0.016644691189821192163, 0.013876128823070747999, 0.011896709945891770095, 0.010411265261972096497,
    0.0092554621827127329177, 0.0083305634333628712565, 0.0075736754879518407950, 0.0069428401072095298657,
    0.0064089941880042070684, 0.0059513701127588477356, 0.0055547335519628013710};

__device__ __forceinline__ double gz_stirling_err(double k) {
    if (k < 16.0) return gz_stirlerr[(int)k];
    const double r = 1.0 / k, r2 = r * r;
    return r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 * (1.0 / 1680.0 - r2 * (1.0 / 1188.0)))));
}

// k ln(k / lam) + lam - k without cancellation: the series in v = (k - lam) / (k + lam)
// near k = lam (C. Loader, "Fast and accurate computation of binomial probabilities", 2000).
__device__ __forceinline__ double gz_bd0(double k, double lam) {
    if (fabs(k - lam) < 0.1 * (k + lam)) {
        const double v = (k - lam) / (k + lam), v2 = v * v;
        double s = (k - lam) * v, ej = 2.0 * k * v;
        for (int j = 1; j < 40; j++) {
            ej *= v2;
            const double s1 = s + ej / (double)(2 * j + 1);
            if (s1 == s) break;
            s = s1;
        }
        return s;
    }
    return k * log(k / lam) + lam - k;
}

// Poisson pmf with relative error ~1e-15 at any lam (the saddle-point form of the same
// paper; exp(k ln lam - lam - ln k!) would lose ~|k ln lam| ulp to cancellation, 1e-10 at
// lam ~ 1e5).
__device__ __forceinline__ double gz_pmf_acc(double lam, double k) {
    if (k == 0.0) return exp(-lam);
    return exp(-gz_stirling_err(k) - gz_bd0(k, lam)) * rsqrt(2.0 * M_PI * k);
}

// Estimated draw time of a tile in units of 1 ns: about 0.7 us per dense record, 1.5 ns per
// table entry and 1.6 us per expected sparse point (a least-squares fit to the draw
// workgroups' times at 91x81x101, F=40, R=500, R^2 0.94); the dispatch plan sorts by it.
__device__ __forceinline__ double gz_tile_cost(int nd, int ntab, double lam_sparse) {
    return 700.0 * nd + 1.5 * ntab + 1600.0 * lam_sparse;
}

// ---------------------------------------------------------------------------
// prep: classify and compact one tile of one frequency
// ---------------------------------------------------------------------------
constexpr int GZ_PREP_FG = 8;   // frequencies per prep workgroup
__global__ void __launch_bounds__(GZ_TILE) k_gz_prep(GapArgs A) {
    __shared__ int s_wd[GZ_TILE / 64], s_ws[GZ_TILE / 64], s_wn[GZ_TILE / 64], s_wb[GZ_TILE / 64];
    __shared__ double s_wsum[GZ_TILE / 64];
    __shared__ double s_plam[GZ_TILE];   // dense records' lambda and first count, compacted: the
    __shared__ int s_pkb[GZ_TILE];       // pmf anchors are evaluated with full waves
    // workgroup per (tile, group of GZ_PREP_FG frequencies): each thread reads its bin's
    // GZ_PREP_FG consecutive values of number and h2fdf (whole cache lines, instead of one
    // 8-byte word per line and frequency), then the frequencies are classified in turn
    // XCD-aware order: workgroup w runs on XCD w % 8, and the workgroups of one XCD take the
    // frequency groups of one tile in turn, so that the 128-byte lines a tile's rows share
    // between groups are fetched once into that XCD's L2
    const int ngroups = (A.F + GZ_PREP_FG - 1) / GZ_PREP_FG;
    const int w = blockIdx.x;
    const int k8 = w / GZ_XCDS;
    const int tile = (k8 / ngroups) * GZ_XCDS + (w % GZ_XCDS);
    if (tile >= A.ntiles) return;
    const int f0 = (k8 % ngroups) * GZ_PREP_FG;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int F = A.F;
    const int64_t b = (int64_t)tile * GZ_TILE + tid;
    const bool skip_zero_h = (A.flags & HD_FLAG_SKIP_ZERO_H) != 0;
    const int nf = min(GZ_PREP_FG, F - f0);
    double v_lam[GZ_PREP_FG], v_h[GZ_PREP_FG];
#pragma unroll
    for (int k = 0; k < GZ_PREP_FG; k++) {
        const bool ok = b < A.B && k < nf;
        v_lam[k] = ok ? A.number[b * F + f0 + k] : 0.0;
        v_h[k] = ok ? A.h2fdf[b * F + f0 + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < GZ_PREP_FG; k++) {
        if (k >= nf) break;
        const int f = f0 + k;
        if (k > 0) __syncthreads();   // the previous frequency's readers of s_w* are done
        double lam = 0.0, h = 0.0;
        bool in = b < A.B;
        if (in) {
            lam = v_lam[k];
            h = v_h[k];
            if (!(lam > 0.0)) in = false;
            if (A.mode == HD_SEL_SORTED && skip_zero_h && h == 0.0) in = false;
            // argmax: h <= 0 adds n*h = 0 and can never be selected
            if (A.mode == HD_SEL_ARGMAX && !(h > 0.0)) in = false;
        }
        const bool dense = in && (lam >= A.tau);
        const bool sparse = in && !dense;
        int kb = 0, kt = 0, nwin = 0;
        if (dense && lam <= A.thresh) gz_window(lam, A.cap, kb, kt, nwin);
        const uint64_t md = __ballot(dense), ms = __ballot(sparse), mb = __ballot(dense && nwin == 0);
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        double x = sparse ? lam : 0.0;
        int xn = nwin;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double y = __shfl_up(x, off, 64);
            const int yn = __shfl_up(xn, off, 64);
            if (lane >= off) {
                x += y;
                xn += yn;
            }
        }
        if (lane == 63) {
            s_wsum[wave] = x;
            s_wn[wave] = xn;
        }
        if (lane == 0) {
            s_wd[wave] = __popcll(md);
            s_ws[wave] = __popcll(ms);
            s_wb[wave] = __popcll(mb);
        }
        __syncthreads();
        int od = 0, os = 0, on = 0, nd = 0, ns = 0, ntot = 0, nbig = 0;
        double osum = 0.0;
        for (int w = 0; w < GZ_TILE / 64; w++) {
            if (w < wave) {
                od += s_wd[w];
                os += s_ws[w];
                on += s_wn[w];
                osum += s_wsum[w];
            }
            nd += s_wd[w];
            ns += s_ws[w];
            ntot += s_wn[w];
            nbig += s_wb[w];
        }
        const int64_t ft = (int64_t)f * A.ntiles + tile;
        const int64_t reg = ft * GZ_TILE;
        const int64_t rreg = ft * GZ_REC;
        const uint32_t rk = (A.mode == HD_SEL_SORTED && in) ? (uint32_t)A.rank[b] : 0u;
        const uint64_t word = ((uint64_t)(uint32_t)b << 32) | rk;
        if (dense) {
            const int i = od + __popcll(md & lt);
            DRec r;
            r.lam = lam;
            r.h = h;
            r.word = word;
            r.kb = kb;
            r.kt = kt;
            r.n = nwin;
            r.p0 = 0.0;               // pmf anchor p(kb): below, with the records compacted
            r.off = on + xn - nwin;   // exclusive prefix of the table lengths
            s_plam[i] = lam;
            s_pkb[i] = (nwin > 0) ? kb : -1;
            A.drec[rreg + i] = r;
            A.rec_end[rreg + i] = r.off + nwin;
            for (int k = 0; k < A.npar; k++)
                A.d_par[(int64_t)k * F * A.ntiles * GZ_TILE + reg + i] = A.par[((int64_t)k * A.B + b) * F + f];
        }
        if (sparse) {
            const int i = os + __popcll(ms & lt);
            const double cum = osum + x;
            A.s_cum[reg + i] = cum;
            A.s_h[reg + i] = h;
            A.s_word[reg + i] = word;
            for (int k = 0; k < A.npar; k++)
                A.sp_par[(int64_t)k * F * A.ntiles * GZ_TILE + reg + i] = A.par[((int64_t)k * A.B + b) * F + f];
            if (i == ns - 1) {   // the last sparse bin writes the pseudo-record for the tile's point count
                DRec r;
                r.lam = cum;
                r.h = 0.0;
                r.word = 0;
                gz_window(cum, A.cap, r.kb, r.kt, r.n);
                r.p0 = (r.n > 0) ? gz_pmf_acc(cum, (double)r.kb) : 0.0;
                r.off = ntot;
                A.drec[rreg + nd] = r;
                A.rec_end[rreg + nd] = ntot + r.n;
            }
        }
        if (tid == 0) {
            A.cnt[ft * 3] = nd;
            A.cnt[ft * 3 + 1] = ns;
            A.cnt[ft * 3 + 2] = nbig;
            double lam_sparse = 0.0;
            for (int w = 0; w < GZ_TILE / 64; w++) lam_sparse += s_wsum[w];
            A.tile_cost[ft] = (float)gz_tile_cost(nd, ntot, lam_sparse);
        }
        __syncthreads();   // the compacted (lambda, kb) of every dense record are in LDS
        if (tid < nd && s_pkb[tid] >= 0) A.drec[rreg + tid].p0 = gz_pmf_acc(s_plam[tid], (double)s_pkb[tid]);
    }
}

// Recording hook: the count drawn for (bin of `word`, f, r).  A bin is drawn by one thread per
// (f, r), so plain stores suffice; sparse points add one each.
__device__ __forceinline__ void gz_record(const GapArgs& A, uint64_t word, int f, int r, double n, bool add) {
    double* p = A.counts_out + ((int64_t)(word >> 32) * A.F + f) * A.R + r;
    *p = add ? *p + n : n;
}

__global__ void k_gz_zero(double* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// plan: the draw workgroups' dispatch order
// ---------------------------------------------------------------------------
// A draw workgroup's time follows its (chunk, f) item's work (the sum of its
// tiles' gz_tile_cost), and the items differ by 10x.  Dispatched in (chunk, f)
// order, long items that start late leave most of the GPU idle at the end, so
// the items go out by decreasing cost instead: a counting sort into buckets of
// 2^(1/4) in cost (the longest-processing-time-first rule, to within that
// ratio).  The order within a bucket follows atomics and may vary between
// calls; results do not depend on it (each item writes its own partials from
// streams keyed by tile, f and realization).
constexpr int GZ_PLAN_BUCKETS = 128;
constexpr int GZ_PLAN_THREADS = 1024;
__global__ void __launch_bounds__(GZ_PLAN_THREADS) k_gz_plan(GapArgs A) {
    __shared__ int s_off[GZ_PLAN_BUCKETS];
    const int tid = threadIdx.x;
    const int n = A.nchunks * A.F;
    for (int i = tid; i < GZ_PLAN_BUCKETS; i += GZ_PLAN_THREADS) s_off[i] = 0;
    __syncthreads();
    for (int it = tid; it < n; it += GZ_PLAN_THREADS) {
        const int f = it / A.nchunks, chunk = it - f * A.nchunks;
        const int t0 = chunk * GZ_TILES_PER_CHUNK;
        const float* tc = A.tile_cost + (int64_t)f * A.ntiles;
        double cost = 1.0;
#pragma unroll
        for (int k = 0; k < GZ_TILES_PER_CHUNK; k++)   // contiguous, every load issued first
            cost += (t0 + k < A.ntiles) ? (double)tc[min(t0 + k, A.ntiles - 1)] : 0.0;
        const int b = (int)fmin(fmax((double)(GZ_PLAN_BUCKETS - 1) - floor(4.0 * log2(cost)), 0.0),
                                (double)(GZ_PLAN_BUCKETS - 1));
        A.plan_b[it] = b;
        atomicAdd(&s_off[b], 1);
    }
    __syncthreads();
    if (tid == 0) {   // exclusive scan of the bucket sizes (heaviest bucket first)
        int sum = 0;
        for (int i = 0; i < GZ_PLAN_BUCKETS; i++) {
            const int c = s_off[i];
            s_off[i] = sum;
            sum += c;
        }
    }
    __syncthreads();
    for (int it = tid; it < n; it += GZ_PLAN_THREADS) A.plan[atomicAdd(&s_off[A.plan_b[it]], 1)] = it;
}

// ---------------------------------------------------------------------------
// draw
// ---------------------------------------------------------------------------
// Per-realization candidate list of the LM best instances; the worst of the
// first L slots is cached.  The sorted walk ranks instances by their position
// in the order (32-bit key; the bin is recovered from the order at the end);
// argmax ranks by (h descending, bin ascending).
// LM = 1, 5, 10 are instantiated for exactly L = LM (the slot of the worst
// kept instance is then a compile-time constant); LM = 4, 16 serve any L up to
// LM with a run-time L.
__host__ __device__ constexpr bool gz_exact(int LM) { return LM != 4 && LM != 16; }
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
        if (gz_exact(LM)) L = LM;
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
    __device__ __forceinline__ uint32_t k32(int j) const { return k[j]; }   // 0xffffffff: none
    // the list as stored by gz_store (32-bit ranks)
    __device__ __forceinline__ void load(const uint64_t* pk, const int32_t*, int L, int64_t stride) {
        if (gz_exact(LM)) L = LM;
        const uint32_t* p = (const uint32_t*)pk;
#pragma unroll
        for (int i = 0; i < LM; i++) {
            k[i] = (i < L) ? p[i * stride] : 0xffffffffu;
            if (i == L - 1) wk = k[i];
        }
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
        if (gz_exact(LM)) L = LM;
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
    __device__ __forceinline__ void load(const uint64_t* pk, const int32_t* pb, int L, int64_t stride) {
        if (gz_exact(LM)) L = LM;
#pragma unroll
        for (int i = 0; i < LM; i++) {
            k[i] = (i < L) ? pk[i * stride] : GZ_KEY_NONE;
            b[i] = (i < L) ? pb[i * stride] : 0x7fffffff;
            if (i == L - 1) {
                wk = k[i];
                wb = b[i];
            }
        }
    }
    __device__ __forceinline__ int32_t bin_at(int j, const int64_t*) const { return b[j]; }
};

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
    __device__ __forceinline__ void add(double n, double h, uint64_t word, const double* p, int64_t pstride, int npar,
                                        int L) {
        if (MODE == HD_SEL_SORTED ? !(n >= 1.0) : (n == 0.0)) return;
        tot += n * h;
#pragma unroll
        for (int k = 0; k < NP; k++)
            if (k < npar) pt[k] += n * h * p[k * pstride];
        if (MODE != HD_SEL_NONE && n > 0 && list.admits(word, h)) {
            const int ninst = (int)fmin(ceil(n), (double)L);
            for (int i = 0; i < ninst && list.admits(word, h); i++) list.insert(word, h, L);
        }
    }
};

// 1/x for x >= 1: hardware reciprocal refined by two Newton steps (about an
// ulp, against ~10 instructions for the correctly rounded divide).
__device__ __forceinline__ double gz_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// CDF value c in [0, 1] as a 64-bit fixed-point fraction floor(c 2^64) split
// into its high and low 32-bit words (c >= 1 saturates); exact.
__device__ __forceinline__ void gz_fixed64(double c, uint32_t& h, uint32_t& l) {
    if (!(c < 1.0)) {
        h = l = 0xFFFFFFFFu;
        return;
    }
    const double t = c * 4294967296.0;   // exact
    h = (uint32_t)t;
    l = (uint32_t)((t - (double)h) * 4294967296.0);   // the fraction is exact, so is its scaling
}

// Inclusive scans of doubles over groups of G lanes (G = 16: each DPP row; G = 64:
// the wave) by DPP moves: row shifts by 1, 2, 4, 8, then for G = 64 the broadcasts
// of lane 15 and lane 31 to the rows after them; in registers instead of LDS
// permutes.  A lane whose source lies outside its group receives the operation's
// identity through the DPP "old" operand (rows outside the row mask keep it too),
// so no selects are needed.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double gz_dpp_id(double v, double id) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(id), __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(id), __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int G, class Op>
__device__ __forceinline__ double gz_group_scan(double v, double id, Op op) {
    v = op(gz_dpp_id<0x111>(v, id), v);   // row_shr:1
    v = op(gz_dpp_id<0x112>(v, id), v);   // row_shr:2
    v = op(gz_dpp_id<0x114>(v, id), v);   // row_shr:4
    v = op(gz_dpp_id<0x118>(v, id), v);   // row_shr:8
    if (G == 64) {
        v = op(gz_dpp_id<0x142, 0xA>(v, id), v);   // row_bcast:15 into rows 1 and 3
        v = op(gz_dpp_id<0x143, 0xC>(v, id), v);   // row_bcast:31 into rows 2 and 3
    }
    return v;
}

// the value of the previous lane of the group (the identity on the group's first lane)
template <int G>
__device__ __forceinline__ double gz_group_prev(double v, double id) {
    return (G == 64) ? gz_dpp_id<0x138>(v, id) : gz_dpp_id<0x111>(v, id);   // wave_shr:1, row_shr:1
}

struct GzMul { __device__ double operator()(double a, double b) const { return a * b; } };
struct GzAdd { __device__ double operator()(double a, double b) const { return a + b; } };
struct GzMax { __device__ double operator()(double a, double b) const { return fmax(a, b); } };

// Build the CDF table of a record with a group of G lanes (G = 16: four records
// per wave, one per DPP row; G = 64: one record per wave), as 64-bit fixed point:
// hi[0, n) the high words, hi[n, 2n) the low words, entry j = floor(2^64 CDF(kt + j))
// with the CDF summed from kb.  The counts kb .. kt + n - 1 are split into
// contiguous segments of the group's lanes; p(kb) = p0 comes from the prep kernel
// and the pmf follows p(k+1) = p(k) lam / (k+1).  Pass 1: each lane's pmf relative
// to its segment's first count, their product and their sum (the running relative
// sums at the table's counts parked in the table's own words); a group prefix
// product gives each segment's first pmf, a group prefix sum of the segment sums
// the carries.  Pass 2 writes carry + first pmf * relative sum (one rounding, fma)
// for the counts from kt on (lanes wholly below kt write nothing: the lower tail is
// only summed).  (Relative error of the CDF about
// (kt + n - kb) ulp, then rounding down to a multiple of 2^-64.)  Four records per
// wave share the fixed cost of the scans; long tables take a whole wave, which
// shortens each lane's serial recurrence (the builds of a batch are its critical
// path).  Arguments are the group's record (n = 0 and kt = kb: no table, the group
// idles); gl = lane index within the group; the whole wave calls (the scans read
// every lane).
typedef __attribute__((address_space(3))) uint32_t gz_lds_u32;
template <int G>
__device__ __noinline__ void gz_build(gz_lds_u32* hi, double lam, int kb, int kt, int n, double p0, int gl) {
    gz_lds_u32* lo = hi + n;
    const int nb = kt + n - kb;          // counts summed
    const int s = (nb + G - 1) / G;
    const int j0 = gl * s;
    const int m = min(max(nb - j0, 0), s);
    const double kf = (double)(kb + j0);
    const int jt = kt - kb - j0;         // segment index of the table's first count
    // pass 1; the relative sum at every count of the table is parked in that entry's own two
    // words (low half in the high-word slot, high half in the low-word slot) for pass 2
    double prod = 1.0, rel = 0.0;
    for (int j = 0; j < m; j++) {
        rel += prod;
        if (j >= jt) {
            hi[j0 + j - (kt - kb)] = (uint32_t)__double2loint(rel);
            lo[j0 + j - (kt - kb)] = (uint32_t)__double2hiint(rel);
        }
        prod *= lam * gz_rcp(kf + (double)(j + 1));
    }
    const double pre = gz_group_prev<G>(gz_group_scan<G>(prod, 1.0, GzMul()), 1.0);   // over the lanes before
    const double pstart = p0 * pre;      // p(kb + j0)
    const double local = pstart * rel;
    const double carry = gz_group_prev<G>(gz_group_scan<G>(local, 0.0, GzAdd()), 0.0);   // sum of those before
    const double last = (m > 0) ? __fma_rn(pstart, rel, carry) : -1.0;
    // The carries are rounded independently of the segment sums, so two
    // neighbouring segments can meet out of order by an ulp; raise each
    // segment to the largest value before it, which makes the table exactly
    // non-decreasing (so the inversion is a monotone map of the uniform).
    const double ninf = -__builtin_inf();
    const double before = gz_group_prev<G>(gz_group_scan<G>(last, ninf, GzMax()), ninf);
    // pass 2, for the counts of the table: the parked relative sums (each lane reads back only
    // what it wrote) and the same fma, so a segment's last value is exactly `last`
    for (int j = max(jt, 0); j < m; j++) {
        const int e = j0 + j - (kt - kb);
        const double r = __hiloint2double((int)lo[e], (int)hi[e]);
        uint32_t h, l;
        gz_fixed64(fmax(__fma_rn(pstart, r, carry), before), h, l);
        hi[e] = h;
        lo[e] = l;
    }
}

// Counts summed (kt + n - kb) up to which a table is built by a 16-lane row, four per wave;
// longer tables take a whole wave.
#ifndef HD_GZ_ROW_MAX
#define HD_GZ_ROW_MAX 256
#endif

// A short table built by one thread in one pass: the same entries as gz_build
// (the CDF summed from kb by the pmf recurrence, floor(2^64 CDF) as high and low
// words), the sum running in order, so the entries are non-decreasing as they
// are.  For tables whose counts kb .. kt + n - 1 number at most GZ_SEQ_MAX: a
// thread per table avoids the wave scans, which cost more than such a table.
#ifndef HD_GZ_SEQ_MAX
#define HD_GZ_SEQ_MAX 32
#endif
constexpr int GZ_SEQ_MAX = HD_GZ_SEQ_MAX;
__device__ __forceinline__ void gz_build_seq(uint32_t* hi, double lam, int kb, int kt, int n, double p0) {
    uint32_t* lo = hi + n;
    double p = p0, c = 0.0;
    int k = kb;
    for (; k < kt; k++) {   // the lower tail, summed into entry 0
        c += p;
        p *= lam * gz_rcp((double)(k + 1));
    }
    for (int j = 0; j < n; j++, k++) {
        c += p;
        uint32_t h, l;
        gz_fixed64(c, h, l);
        hi[j] = h;
        lo[j] = l;
        p *= lam * gz_rcp((double)(k + 1));
    }
}

// Inversion by the stream: the smallest count k with CDF(k) > u, u = (x, w) / 2^64
// for the next words x, w.  A branch-free binary search over the table's high
// words finds the first entry j with hi[j] >= x; when that high word equals x the
// low words decide, so only then is w taken (probability about n 2^-32 per
// draw).  j runs 0 .. n (n: u at or above the last entry); j = 0 with a summed
// lower tail (kt > kb) or j = n go to the tail walk (which takes w if the search
// did not), else the count is kt + j.  hi[n] (the first low word) may be read by
// the search but is not used.  (The candidate is carried as an address, so a
// step costs one add, one compare and one select, on 4-byte words.)  Two searches
// run interleaved (gz_draw2) so their LDS round trips overlap.
struct GzSearch {
    const gz_lds_u32* hi;
    const gz_lds_u32* p;   // invariant: the first j with hi[j] >= x is in [p - hi, p - hi + len]
    int n, len;
    uint32_t x;
};

// Entry index of an LDS word address (32-bit address arithmetic).
__device__ __forceinline__ int gz_idx(const gz_lds_u32* p, const gz_lds_u32* base) {
    return (int)(((uint32_t)(size_t)p - (uint32_t)(size_t)base) >> 2);
}

__device__ __forceinline__ void gz_search_start(GzSearch& s, const gz_lds_u32* hi, int n, uint32_t x) {
    s.hi = hi;
    s.p = hi;
    s.n = n;
    s.len = n;
    s.x = x;
}

__device__ __forceinline__ void gz_search_step(GzSearch& s) {
    const int half = s.len >> 1;
    const gz_lds_u32* q = s.p + half;
    s.p = (*q < s.x) ? q : s.p;
    s.len -= half;
}

// First entry j (0 .. n) with (hi[j], lo[j]) > (x, w); w is taken only on a tie.
template <class Words>
__device__ __forceinline__ int gz_search_finish(const GzSearch& s, Words& rs, uint32_t& w, bool& has_w) {
    const uint32_t h0 = s.p[0], h1 = s.p[1];
    const bool up = h0 < s.x;
    int j = gz_idx(s.p, s.hi) + (up ? 1 : 0);
    has_w = false;
    if (j < s.n && (up ? h1 : h0) == s.x) {
        w = rs.next();
        has_w = true;
        const gz_lds_u32* lo = s.hi + s.n;
        while (j < s.n && s.hi[j] == s.x && lo[j] <= w) j++;
    }
    return j;
}

// The count for u = (x, w) / 2^64 outside the table (j = 0 below its first entry,
// j = n at or above its last), by the pmf recurrence from the table's edge in
// units of 2^-64: downwards p(k-1) = p(k) k / lam to the support's first count
// kb, upwards p(k+1) = p(k) lam / (k+1) to its last, khi (the CDF values there
// differ from u by less than 2^-52 of the tail mass left, which is below 2^-12
// of the whole).  Each edge pmf is the saddle-point value (relative error
// ~1e-15).
__device__ __noinline__ int gz_tail(const gz_lds_u32* hi, int n, int j, uint32_t x, uint32_t w, double lam, int kb,
                                    int kt) {
    const gz_lds_u32* lo = hi + n;
    const uint64_t u = ((uint64_t)x << 32) | w;
    constexpr double two64 = 18446744073709551616.0;
    if (j == 0) {   // CDF(kt) > u: the largest k with CDF(k - 1) <= u
        const uint64_t f0 = ((uint64_t)hi[0] << 32) | lo[0];
        double d = (double)(f0 - u);                        // CDF(k) - u, k = kt
        double p = gz_pmf_acc(lam, (double)kt) * two64;     // p(k)
        int k = kt;
        while (k > kb && d > p) {
            d -= p;
            p *= (double)k / lam;
            k--;
        }
        return k;
    }
    const int kc = kt + n - 1;   // CDF(kc) <= u: the first k > kc with CDF(k) > u
    double slo, shi;
    gz_support(lam, slo, shi);
    const int khi = (int)shi;
    if (kc >= khi) return kc;    // above the support's CDF only by rounding
    const uint64_t fl = ((uint64_t)hi[n - 1] << 32) | lo[n - 1];
    double d = (double)(u - fl);                             // u - CDF(k - 1), k = kc + 1
    int k = kc + 1;
    double p = gz_pmf_acc(lam, (double)k) * two64;           // p(k)
    while (k < khi && d >= p) {
        d -= p;
        k++;
        p *= lam / (double)k;
    }
    return k;
}

// A table as the draws see it: its words, length, first counts, and the staged
// header (lam is read from there only by a tail walk).
struct GzTab {
    const gz_lds_u32* hi;
    int n, kb, kt;
    const BRec* rec;
};

template <class Words>
__device__ __forceinline__ int gz_count(const GzTab& t, int j, uint32_t x, uint32_t w, bool has_w, Words& rs) {
    if ((j == 0 && t.kt > t.kb) || j == t.n) {
        if (!has_w) w = rs.next();
        return gz_tail(t.hi, t.n, j, x, w, t.rec->lam, t.kb, t.kt);
    }
    return t.kt + j;
}

// One draw from a table (count kt + j of its entries, tails beyond).
template <class Words>
__device__ __forceinline__ int gz_draw1(const GzTab& t, Words& rs) {
    GzSearch s;
    const uint32_t x = rs.next();
    gz_search_start(s, t.hi, t.n, x);
    while (s.len > 1) gz_search_step(s);
    {   // the common case, as in gz_draw2
        const uint32_t h0 = s.p[0], h1 = s.p[1];
        const bool up = h0 < x;
        const int j = gz_idx(s.p, s.hi) + (up ? 1 : 0);
        // fast: no tie on the high word and j in [lo, n), lo = 1 when entry 0 carries a summed
        // lower tail (kt > kb), else 0 -- one unsigned compare for the range
        const int lo = (t.kt > t.kb) ? 1 : 0;
        const bool fast = ((uint32_t)(j - lo) < (uint32_t)(t.n - lo)) & ((up ? h1 : h0) != x);
        if (__builtin_amdgcn_ballot_w64(!fast) == 0ull) return t.kt + j;
    }
    uint32_t w = 0;
    bool hw;
    const int j = gz_search_finish(s, rs, w, hw);
    return gz_count(t, j, x, w, hw, rs);
}

// Two draws from tables a and b (n = 0: no draw, count 0).  Words in the order
// x_a, x_b, then w_a, w_b where a tie needs them, then w_a, w_b where a tail
// walk needs them and the search did not take them.

template <class Words>
__device__ __forceinline__ void gz_draw2(const GzTab& ta, const GzTab& tb, Words& rs, int& ca, int& cb) {
    GzSearch a, b;
    const uint32_t xa = (ta.n > 0) ? rs.next() : 0u;
    const uint32_t xb = (tb.n > 0) ? rs.next() : 0u;
    gz_search_start(a, ta.hi, ta.n, xa);
    gz_search_start(b, tb.hi, tb.n, xb);
    while (a.len > 1 && b.len > 1) {
        // both loads issued before either compare, so their LDS latencies overlap
        const int ha = a.len >> 1, hb = b.len >> 1;
        const gz_lds_u32* qa = a.p + ha;
        const gz_lds_u32* qb = b.p + hb;
        const uint32_t va = *qa;
        const uint32_t vb = *qb;
        a.p = (va < a.x) ? qa : a.p;
        b.p = (vb < b.x) ? qb : b.p;
        a.len -= ha;
        b.len -= hb;
    }
    while (a.len > 1) gz_search_step(a);
    while (b.len > 1) gz_search_step(b);
    {   // the common case: no lane of the wave has a tie on a high word (the low words needed)
        // or a draw outside a table, so every count is kt + j from the candidates alone (one
        // wave-wide test instead of per-lane branches); otherwise the exact path below
        const uint32_t a0 = a.p[0], a1 = a.p[1], b0 = b.p[0], b1 = b.p[1];
        const bool ua = a0 < xa, ub = b0 < xb;
        const int ja = gz_idx(a.p, a.hi) + (ua ? 1 : 0), jb = gz_idx(b.p, b.hi) + (ub ? 1 : 0);
        // fast: no tie on a high word and j in [lo, n) (lo = 1 when entry 0 carries a summed lower
        // tail), as in gz_draw1; straight-line code on masks
        const int loa = (ta.kt > ta.kb) ? 1 : 0, lob = (tb.kt > tb.kb) ? 1 : 0;
        const bool fa = ((uint32_t)(ja - loa) < (uint32_t)(ta.n - loa)) & ((ua ? a1 : a0) != xa);
        const bool fb = ((uint32_t)(jb - lob) < (uint32_t)(tb.n - lob)) & ((ub ? b1 : b0) != xb);
        if (__builtin_amdgcn_ballot_w64(((ta.n > 0) & !fa) | ((tb.n > 0) & !fb)) == 0ull) {
            ca = (ta.n > 0) ? ta.kt + ja : 0;
            cb = (tb.n > 0) ? tb.kt + jb : 0;
            return;
        }
    }
    uint32_t wa = 0, wb = 0;
    bool hwa = false, hwb = false;
    const int ja = (ta.n > 0) ? gz_search_finish(a, rs, wa, hwa) : 1;
    const int jb = (tb.n > 0) ? gz_search_finish(b, rs, wb, hwb) : 1;
    ca = (ta.n > 0) ? gz_count(ta, ja, xa, wa, hwa, rs) : 0;
    cb = (tb.n > 0) ? gz_count(tb, jb, xb, wb, hwb, rs) : 0;
}

// The table of a staged header (the header is the same for the whole wave: its
// fields go to scalar registers); present = false gives an empty table.
__device__ __forceinline__ GzTab gz_tab(const uint32_t* tab, const BRec* b, bool present) {
    GzTab t;
    t.n = present ? __builtin_amdgcn_readfirstlane(b->n) : 0;
    t.hi = (const gz_lds_u32*)(tab + 2 * __builtin_amdgcn_readfirstlane(b->off));
    t.kb = __builtin_amdgcn_readfirstlane(b->kb);
    t.kt = __builtin_amdgcn_readfirstlane(b->kt);
    t.rec = b;
    return t;
}

// The table of a record read from global memory through the constant address space (the
// fields are wave-uniform: scalar loads); off is relative to the batch's first table.
typedef const __attribute__((address_space(4))) DRec* gz_cdrec;
__device__ __forceinline__ DRec gz_load_rec(gz_cdrec p) {
    DRec d;
    d.lam = p->lam;
    d.h = p->h;
    d.word = p->word;
    d.p0 = p->p0;
    d.kb = p->kb;
    d.n = p->n;
    d.off = p->off;
    d.kt = p->kt;
    return d;
}
__device__ __forceinline__ GzTab gz_tab_rec(const uint32_t* tab, const DRec& d, int base, const BRec* staged,
                                            bool present) {
    GzTab t;
    t.n = present ? d.n : 0;
    t.hi = (const gz_lds_u32*)(tab + 2 * (d.off - base));
    t.kb = d.kb;
    t.kt = d.kt;
    t.rec = staged;
    return t;
}

// Counts without a table: Normal above the threshold, inversion below 10,
// PTRS otherwise; stream (a, f, r, tag).
__device__ __forceinline__ double gz_fallback(uint32_t a, uint32_t f, uint32_t r, uint32_t tag, uint64_t seed,
                                              double lam, double thresh) {
    Stream st;
    st.init(a, f, r, tag, seed);
    if (lam > thresh) return normal_count(st, lam);
    if (lam < 10.0) return poisson_inversion(st.next(), lam, -expm1(-lam), exp(-lam));
    return poisson_ptrs(st, lam);
}

template <int MODE, int LM, int NP>
__device__ __forceinline__ void gz_store(const GapArgs& A, const Acc<MODE, LM, NP>& acc, int64_t pr) {
    const int F = A.F, R = A.R, L = gz_exact(LM) ? LM : A.L;
    A.p_total[pr] = acc.tot;
#pragma unroll
    for (int k = 0; k < NP; k++)
        if (k < A.npar) A.p_ptot[(int64_t)k * F * A.nchunks * R + pr] = acc.pt[k];
    if (MODE != HD_SEL_NONE) {
#pragma unroll
        for (int j = 0; j < LM; j++) {
            if (j >= L) continue;
            const int64_t o = ((pr / R) * L + j) * R + (pr % R);
            if constexpr (MODE == HD_SEL_SORTED) {   // ranks only (the merge takes bins from the order)
                ((uint32_t*)A.p_key)[o] = acc.list.k32(j);
            } else {
                A.p_key[o] = acc.list.key_at(j);
                A.p_bin[o] = acc.list.bin_at(j, A.order);
            }
        }
    }
}

// REC: also write every drawn count to A.counts_out (the recording test hook); the production
// instantiations have REC = false and carry no trace of it.
template <int MODE, int LM, int NP, bool REC>
// (6 waves per SIMD: with the LDS below three 512-thread workgroups share a CU)
__global__ void __launch_bounds__(GZ_MAXBS, 6) k_gz_draw(GapArgs A) {
    extern __shared__ double gz_lds[];
    __shared__ int s_next;             // next build item of the batch
    __shared__ int s_nlarge;           // long tables of the batch (a wave each)
    __shared__ int s_nrow;             // mid-size tables (a 16-lane row each)
    __shared__ int s_nshort;           // short tables (a lane each)
    __shared__ int s_large[GZ_BATCH_REC];
    __shared__ int s_row[GZ_BATCH_REC];
    __shared__ int s_short[GZ_BATCH_REC];
    uint32_t* tab = (uint32_t*)gz_lds;   // A.cap 8-byte slots: per record n high words, then n low words
    double* scum = gz_lds + A.cap;     // GZ_TILE doubles
    BRec* brec = (BRec*)(scum + GZ_TILE);   // GZ_BATCH_REC staged record headers
    const int item = A.plan[blockIdx.x];   // heaviest items first
    const int f = item / A.nchunks;
    const int chunk = item - f * A.nchunks;
    const int tid = threadIdx.x, lane = tid & 63;
    const int bs = blockDim.x;
    const int F = A.F, R = A.R, L = A.L;
    const int t0 = chunk * GZ_TILES_PER_CHUNK;
    const int t1 = min(t0 + GZ_TILES_PER_CHUNK, A.ntiles);
    const int64_t pstride = (int64_t)F * A.ntiles * GZ_TILE;
    const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);

    for (int rbase = 0; rbase < R; rbase += bs) {
        const int r = rbase + tid;
        const bool active = r < R;
        Acc<MODE, LM, NP> acc;
        acc.init();
        for (int tile = t0; tile < t1; tile++) {
            const int64_t ft = (int64_t)f * A.ntiles + tile;
            const int64_t reg = ft * GZ_TILE;
            const int64_t rreg = ft * GZ_REC;
            const int nd = A.cnt[ft * 3];
            const int ns = A.cnt[ft * 3 + 1];
            const int nrec = nd + (ns > 0 ? 1 : 0);
            // the realization's stream for this tile: records in order take one word each
            // (two when the 32-bit draw is open), then the sparse points two words each
            Xo128 rs;
            rs.seed(philox4x32_10(U4{(uint32_t)tile, (uint32_t)f, (uint32_t)r, 6u << 16}, k0, k1));
            int c = 0;
            while (c < nrec) {
                // batch [c, c1): consecutive records whose windows fit the table
                const int base = (c == 0) ? 0 : A.rec_end[rreg + c - 1];
                const int limit = base + A.cap;
                int c1 = c;
                for (;;) {
                    const int idx = c1 + lane;
                    const int e = (idx < nrec) ? A.rec_end[rreg + idx] : 0x7fffffff;
                    const int adv = __popcll(__ballot(e <= limit));
                    c1 += adv;
                    if (adv < 64) break;
                }
                c1 = max(c1, c + 1);   // every window fits alone (n <= cap); keeps progress unconditional
                c1 = min(c1, c + GZ_BATCH_REC);
                if (tid == 0) {
                    s_next = 0;
                    s_nlarge = 0;
                    s_nrow = 0;
                    s_nshort = 0;
                }
                __syncthreads();   // the previous batch's draws are done with tab / scum
                // one thread per record of the batch stages its header and lists its table as
                // short (at most GZ_SEQ_MAX counts summed) or long
                if (c + tid < c1) {
                    const DRec rec = A.drec[rreg + c + tid];
                    brec[tid] = BRec{rec.h, rec.word, rec.lam, rec.kt, rec.n, rec.off - base, rec.kb, rec.p0};
                    const int nb = rec.kt + rec.n - rec.kb;
                    if (rec.n > 0 && nb <= GZ_SEQ_MAX) s_short[atomicAdd(&s_nshort, 1)] = tid;
                    else if (rec.n > 0 && nb <= HD_GZ_ROW_MAX) s_row[atomicAdd(&s_nrow, 1)] = tid;
                    else if (rec.n > 0) s_large[atomicAdd(&s_nlarge, 1)] = tid;
                }
                __syncthreads();
                // build items handed out to the waves one at a time, the longest serial work
                // first: long tables (a wave each), then mid-size tables four at a time (a
                // 16-lane row each), then groups of 64 short tables (a lane each, one
                // sequential pass), so that every lane a build occupies has work
                {
                    const int nl = s_nlarge, nr = s_nrow, nsh = s_nshort;
                    const int nrg = (nr + 3) >> 2, nsg = (nsh + 63) >> 6;
                    for (;;) {
                        int w = 0;
                        if (lane == 0) w = atomicAdd(&s_next, 1);
                        w = __shfl(w, 0, 64);
                        if (w >= nl + nrg + nsg) break;
                        if (w < nl) {
                            const BRec* rec = brec + s_large[w];
                            gz_build<64>((gz_lds_u32*)(tab + 2 * rec->off), rec->lam, rec->kb, rec->kt, rec->n, rec->p0,
                                         lane);
                        } else if (w < nl + nrg) {
                            const int k = (w - nl) * 4 + (lane >> 4);
                            const bool on = k < nr;
                            const BRec* rec = brec + s_row[on ? k : 0];
                            gz_build<16>((gz_lds_u32*)(tab + 2 * rec->off), rec->lam, on ? rec->kb : 0, on ? rec->kt : 0,
                                         on ? rec->n : 0, rec->p0, lane & 15);
                        } else {
                            const int k = (w - nl - nrg) * 64 + lane;
                            if (k < nsh) {
                                const BRec* rec = brec + s_short[k];
                                gz_build_seq(tab + 2 * rec->off, rec->lam, rec->kb, rec->kt, rec->n, rec->p0);
                            }
                        }
                    }
                }
                if (c1 == nrec && ns > 0)
                    for (int j = tid; j < ns; j += bs) scum[j] = A.s_cum[reg + j];
                __syncthreads();
                if (active) {
                    // dense records of the batch (one control path, so the running list needs no
                    // copies between branches), then the tile's sparse pseudo-record if it is here
                    // (records in pairs, their searches interleaved; n = 0: no table, drawn by k_gz_big)
                    // the records' fields are the same for every lane: read through the constant
                    // address space (scalar loads into SGPRs, no VALU moves); the staged LDS copy
                    // serves the builds and the rare tail walk
                    const gz_cdrec cdr = (gz_cdrec)(A.drec + rreg);
                    const int cd = min(c1, nd);
                    for (int i = c; i < cd; i += 2) {
                        const bool hb = i + 1 < cd;
                        const DRec ra = gz_load_rec(cdr + i);
                        const DRec rb = gz_load_rec(cdr + (hb ? i + 1 : i));
                        const GzTab ta = gz_tab_rec(tab, ra, base, brec + (i - c), true);
                        const GzTab tb = gz_tab_rec(tab, rb, base, brec + (hb ? i + 1 - c : i - c), hb);
                        int ka, kb;
                        gz_draw2(ta, tb, rs, ka, kb);
                        if (ta.n > 0) acc.add((double)ka, ra.h, ra.word, A.d_par + reg + i, pstride, A.npar, L);
                        if (tb.n > 0) acc.add((double)kb, rb.h, rb.word, A.d_par + reg + i + 1, pstride, A.npar, L);
                        if (REC) {
                            if (ta.n > 0) gz_record(A, ra.word, f, r, (double)ka, false);
                            if (tb.n > 0) gz_record(A, rb.word, f, r, (double)kb, false);
                        }
                    }
                    for (int i = max(c, nd); i < c1; i++) {
                        const BRec* rec = brec + (i - c);
                        {
                            // sparse bins: Poisson process on the tile's cumulative-lambda axis; the
                            // point count from its table (it always fits: lambda < 512), the
                            // positions uniform on the axis (53-bit uniforms)
                            const double Lam = scum[ns - 1];
                            const int np = gz_draw1(gz_tab(tab, rec, true), rs);
                            for (int q = 0; q < np; q++) {
                                const uint32_t w0 = rs.next();
                                const double xq = u53(w0, rs.next()) * Lam;
                                int lo = 0, hi = ns - 1;
                                while (lo < hi) {   // first j with scum[j] > xq
                                    const int mid = (lo + hi) >> 1;
                                    if (scum[mid] > xq) hi = mid;
                                    else lo = mid + 1;
                                }
                                acc.add(1.0, A.s_h[reg + lo], A.s_word[reg + lo], A.sp_par + reg + lo, pstride, A.npar,
                                        L);
                                if (REC) gz_record(A, A.s_word[reg + lo], f, r, 1.0, true);
                            }
                        }
                    }
                }
                c = c1;
            }
        }
        if (active) gz_store<MODE, LM, NP>(A, acc, ((int64_t)f * A.nchunks + chunk) * R + r);
    }
}

// Records without a table (lambda above the threshold or a window larger than
// the LDS table): reload the chunk's partials, draw, store them back.
template <int MODE, int LM, int NP, bool REC>
__global__ void __launch_bounds__(256) k_gz_big(GapArgs A) {
    const int f = blockIdx.y;
    const int chunk = blockIdx.x;
    const int F = A.F, R = A.R, L = A.L;
    const int t0 = chunk * GZ_TILES_PER_CHUNK;
    const int t1 = min(t0 + GZ_TILES_PER_CHUNK, A.ntiles);
    const int64_t pstride = (int64_t)F * A.ntiles * GZ_TILE;
    int nbig = 0;
    for (int tile = t0; tile < t1; tile++) nbig += A.cnt[((int64_t)f * A.ntiles + tile) * 3 + 2];
    if (nbig == 0) return;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        const int64_t pr = ((int64_t)f * A.nchunks + chunk) * R + r;
        Acc<MODE, LM, NP> acc;
        acc.tot = A.p_total[pr];
#pragma unroll
        for (int k = 0; k < NP; k++) acc.pt[k] = (k < A.npar) ? A.p_ptot[(int64_t)k * F * A.nchunks * R + pr] : 0.0;
        if (MODE != HD_SEL_NONE) {
            const int64_t o = ((pr / R) * L) * R + (pr % R);
            acc.list.load(MODE == HD_SEL_SORTED ? (const uint64_t*)((const uint32_t*)A.p_key + o) : A.p_key + o,
                          A.p_bin + o, L, R);
        }
        for (int tile = t0; tile < t1; tile++) {
            const int64_t ft = (int64_t)f * A.ntiles + tile;
            if (A.cnt[ft * 3 + 2] == 0) continue;
            const int nd = A.cnt[ft * 3];
            for (int i = 0; i < nd; i++) {
                const DRec rec = A.drec[ft * GZ_REC + i];
                if (rec.n != 0) continue;
                const double n = gz_fallback((uint32_t)(rec.word >> 32), (uint32_t)f, (uint32_t)r, 3u, A.seed, rec.lam,
                                             A.thresh);
                acc.add(n, rec.h, rec.word, A.d_par + ft * GZ_TILE + i, pstride, A.npar, L);
                if (REC) gz_record(A, rec.word, f, r, n, false);
            }
        }
        gz_store<MODE, LM, NP>(A, acc, pr);
    }
}

// ---------------------------------------------------------------------------
// hooks for tests: one CDF window built by one wave; draws from given words
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_cdf_table_hook(double lam, int cap, int* kn, uint64_t* cdf_out) {
    extern __shared__ double hook_lds[];
    uint32_t* tab = (uint32_t*)hook_lds;
    const int lane = threadIdx.x;
    int kb, kt, n;
    gz_window(lam, cap, kb, kt, n);
    if (n > 0) {   // the build the draw kernel uses for this length
        if (kt + n - kb <= GZ_SEQ_MAX) {
            if (lane == 0) gz_build_seq(tab, lam, kb, kt, n, gz_pmf_acc(lam, (double)kb));
        } else if (kt + n - kb <= HD_GZ_ROW_MAX) {   // row 0 builds, rows 1-3 idle
            const bool on = lane < 16;
            gz_build<16>((gz_lds_u32*)tab, lam, on ? kb : 0, on ? kt : 0, on ? n : 0, gz_pmf_acc(lam, (double)kb),
                         lane & 15);
        } else {
            gz_build<64>((gz_lds_u32*)tab, lam, kb, kt, n, gz_pmf_acc(lam, (double)kb), lane);
        }
    }
    __syncthreads();
    for (int j = lane; j < n; j += 64) cdf_out[j] = ((uint64_t)tab[j] << 32) | tab[n + j];
    if (lane == 0) {
        kn[0] = kb;
        kn[1] = kt;
        kn[2] = n;
    }
}

struct GzWords {   // a stream that replays given words
    const uint32_t* w;
    int i;
    __device__ uint32_t next() { return w[i++]; }
};

// Draws from given tables (copied into LDS as the kernels hold them) with given
// words: 2 words per draw, or 4 per pair of draws.
__global__ void __launch_bounds__(256) k_table_draw_hook(const uint32_t* tab, int n, double lam, int kb, int kt,
                                                         const uint32_t* words, int64_t m, int* count, int* used) {
    extern __shared__ double hook_lds[];
    uint32_t* t = (uint32_t*)hook_lds;
    __shared__ BRec hdr;
    for (int j = threadIdx.x; j < 2 * n; j += blockDim.x) t[j] = tab[j];
    if (threadIdx.x == 0) hdr = BRec{0.0, 0, lam, kt, n, 0, kb, 0.0};
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    GzWords ws{words + 2 * i, 0};
    count[i] = gz_draw1(GzTab{(const gz_lds_u32*)t, n, kb, kt, &hdr}, ws);
    used[i] = ws.i;
}

__global__ void __launch_bounds__(256) k_table_draw_pair_hook(const uint32_t* tab, int n, double lam, int kb, int kt,
                                                              const uint32_t* tab2, int n2, double lam2, int kb2,
                                                              int kt2, const uint32_t* words, int64_t m, int* count,
                                                              int* used) {
    extern __shared__ double hook_lds[];
    uint32_t* t = (uint32_t*)hook_lds;
    uint32_t* t2 = t + 2 * n;
    for (int j = threadIdx.x; j < 2 * n; j += blockDim.x) t[j] = tab[j];
    __shared__ BRec hdr[2];
    for (int j = threadIdx.x; j < 2 * n2; j += blockDim.x) t2[j] = tab2[j];
    if (threadIdx.x == 0) {
        hdr[0] = BRec{0.0, 0, lam, kt, n, 0, kb, 0.0};
        hdr[1] = BRec{0.0, 0, lam2, kt2, n2, 0, kb2, 0.0};
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    GzWords ws{words + 4 * i, 0};
    int ca, cb;
    gz_draw2(GzTab{(const gz_lds_u32*)t, n, kb, kt, &hdr[0]}, GzTab{(const gz_lds_u32*)t2, n2, kb2, kt2, &hdr[1]}, ws, ca,
             cb);
    count[2 * i] = ca;
    count[2 * i + 1] = cb;
    used[i] = ws.i;
}

int realize_merge_launch(const double* h2fdf, const double* par, const int64_t* order, int64_t B, int F, int R, int L,
                         int mode, int npar, int nchunks, const double* p_total, const double* p_ptot, const uint64_t* p_key,
                         const int32_t* p_bin, double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar,
                         void* group_ws, hipStream_t st);
size_t realize_group_ws(int F, int R, int L, int npar);
int rank_from_order_launch(const int64_t* order, int64_t B, int32_t* rank, hipStream_t st);

static size_t gz_align(size_t x) { return (x + 255) & ~(size_t)255; }

// Draw workgroups always have GZ_MAXBS threads: with fewer realizations the
// threads beyond R only build tables (eight waves share the builds at any R).
static int gz_block(int) { return GZ_MAXBS; }
// LDS table capacity (8-byte slots) of every draw workgroup: table + sparse
// axis + staged record headers + the build counter stay within 50 KB, so that
// three workgroups fit one CU's 160 KB (allocation granularity included).
// Fixed, so that which bins get a table (and so every draw) does not depend
// on the launch shape.
#ifndef HD_GZ_CAP
#define HD_GZ_CAP 5120
#endif
static int gz_cap(int) { return HD_GZ_CAP; }
// draw workgroup LDS: table + sparse axis + staged headers (+ the build counter); three must fit a
// 160 KB CU (with the 64 KB per-workgroup limit of the allocation), and the hook's table as well
constexpr size_t GZ_DRAW_LDS = (size_t)(HD_GZ_CAP + GZ_TILE) * sizeof(double) + (size_t)GZ_BATCH_REC * sizeof(BRec);
static_assert(GZ_DRAW_LDS + 16 <= 65536, "draw workgroup LDS above 64 KB");
// (the static part: the build counters and the three 128-entry lists of k_gz_draw, rounded up)
static_assert(3 * (GZ_DRAW_LDS + 2048) <= 160 * 1024, "three draw workgroups no longer share a CU");
constexpr int GZ_HOOK_CAP_MAX = 8000;   // hd_poisson_cdf_table: 8-byte slots, 64000 B of LDS
static_assert(GZ_HOOK_CAP_MAX * sizeof(double) <= 65536, "hook table above 64 KB");

static size_t gz_ws(int64_t B, int F, int R, int L, int npar) {
    const int64_t nt = cdiv(B, GZ_TILE);
    const int64_t nch = cdiv(nt, GZ_TILES_PER_CHUNK);
    const size_t slots = (size_t)F * nt * GZ_TILE;
    const size_t rslots = (size_t)F * nt * GZ_REC;
    const size_t pr = (size_t)F * nch * R;
    size_t s = gz_align((size_t)B * 4);                    // rank
    s += gz_align(rslots * sizeof(DRec));                  // dense records
    s += gz_align(rslots * 4);                             // record ends
    s += gz_align((size_t)F * nt * 3 * 4);                 // counts
    s += 3 * gz_align(slots * 8);                          // sparse cum, h, word
    s += 2 * gz_align(slots * 8 * (size_t)npar);           // params
    s += gz_align(pr * 8) + gz_align(pr * 8 * (size_t)npar) + gz_align(pr * 8 * (size_t)L) +
         gz_align(pr * 4 * (size_t)L);
    s += gz_align(realize_group_ws(F, R, L, npar));          // first merge stage
    s += 2 * gz_align((size_t)nch * F * 4);                  // dispatch plan, cost buckets
    s += gz_align((size_t)F * nt * 4);                       // tile costs
    return s + 512;
}

template <int MODE, int LM, int NP, bool REC>
static int gz_launch_k(const GapArgs& A, dim3 grid, int bs, size_t lds, hipStream_t st) {
    {
        KTimer timer("k_gz_draw", st);
        hipLaunchKernelGGL((k_gz_draw<MODE, LM, NP, REC>), dim3(grid.x * grid.y), dim3(bs), lds, st, A);
    }
    int rc = check_launch("k_gz_draw");
    if (rc) return rc;
    hipLaunchKernelGGL((k_gz_big<MODE, LM, NP, REC>), grid, dim3(256), 0, st, A);
    return check_launch("k_gz_big");
}

template <int MODE, int LM, int NP>
static int gz_launch(const GapArgs& A, dim3 grid, int bs, size_t lds, hipStream_t st) {
    if (A.counts_out) return gz_launch_k<MODE, LM, NP, true>(A, grid, bs, lds, st);
    return gz_launch_k<MODE, LM, NP, false>(A, grid, bs, lds, st);
}

template <int MODE, int LM>
static int gz_launch_np(const GapArgs& A, int npar, dim3 grid, int bs, size_t lds, hipStream_t st) {
    if (npar == 0) return gz_launch<MODE, LM, 0>(A, grid, bs, lds, st);
    if (MODE == HD_SEL_ARGMAX && npar <= 3) return gz_launch<MODE, LM, 3>(A, grid, bs, lds, st);
    return gz_launch<MODE, LM, 7>(A, grid, bs, lds, st);
}

}  // namespace hd

using namespace hd;

extern "C" size_t hd_realize_workspace_size(int64_t B, int F, int R, int L, int npar) {
    return gz_ws(B, F, R, L, npar);
}

static int realize_impl(const double* number, const double* h2fdf, const int64_t* order, int64_t B, int F, int R, int L,
                        int mode, int flags, double normal_threshold, uint64_t seed, const double* par, int npar,
                        double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar, double* counts_out,
                        void* workspace, size_t ws_bytes, void* stream) {
    if (B < 1 || F < 1 || R < 1 || !number || !h2fdf || !hc2bg) return set_error(HD_ERR_ARG, "hd_realize: bad arguments");
    if (B > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_realize: B=%lld exceeds 2^31-1", (long long)B);
    if (F > 65535) return set_error(HD_ERR_SIZE, "hd_realize: F=%d > 65535", F);
    if (R > 0x7fffffff / 2) return set_error(HD_ERR_SIZE, "hd_realize: R=%d too large", R);
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
    KTimer timer("hd_realize", st);   // the whole launch sequence, rank to merge
    const int bs = gz_block(R);
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
    A.cap = gz_cap(bs);
    A.thresh = normal_threshold;
    A.tau = fmin(HD_GZ_TAU, normal_threshold);   // bins above the threshold are never sparse
    A.seed = seed;
    A.counts_out = counts_out;
    A.ntiles = (int)cdiv(B, GZ_TILE);
    A.nchunks = (int)cdiv(A.ntiles, GZ_TILES_PER_CHUNK);
    const size_t slots = (size_t)F * A.ntiles * GZ_TILE;
    const size_t rslots = (size_t)F * A.ntiles * GZ_REC;
    const size_t pr = (size_t)F * A.nchunks * R;
    char* p = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    int32_t* rank = (int32_t*)p;  p += gz_align((size_t)B * 4);
    A.drec = (DRec*)p;            p += gz_align(rslots * sizeof(DRec));
    A.rec_end = (int32_t*)p;      p += gz_align(rslots * 4);
    A.cnt = (int32_t*)p;          p += gz_align((size_t)F * A.ntiles * 3 * 4);
    A.s_cum = (double*)p;         p += gz_align(slots * 8);
    A.s_h = (double*)p;           p += gz_align(slots * 8);
    A.s_word = (uint64_t*)p;      p += gz_align(slots * 8);
    A.d_par = (double*)p;         p += gz_align(slots * 8 * (size_t)npar);
    A.sp_par = (double*)p;        p += gz_align(slots * 8 * (size_t)npar);
    A.p_total = (double*)p;       p += gz_align(pr * 8);
    A.p_ptot = (double*)p;        p += gz_align(pr * 8 * (size_t)npar);
    A.p_key = (uint64_t*)p;       p += gz_align(pr * 8 * (size_t)L);
    A.p_bin = (int32_t*)p;        p += gz_align(pr * 4 * (size_t)L);
    void* group_ws = p;           p += gz_align(realize_group_ws(F, R, L, npar));
    A.plan = (int32_t*)p;         p += gz_align((size_t)A.nchunks * F * 4);
    A.plan_b = (int32_t*)p;       p += gz_align((size_t)A.nchunks * F * 4);
    A.tile_cost = (float*)p;
    A.rank = (mode == HD_SEL_SORTED) ? rank : nullptr;
    if (mode == HD_SEL_SORTED) {
        int rc = rank_from_order_launch(order, B, rank, st);
        if (rc) return rc;
    }
    if (counts_out) {
        hipLaunchKernelGGL(k_gz_zero, dim3(2048), dim3(256), 0, st, counts_out, B * (int64_t)F * R);
        int rc = check_launch("k_gz_zero");
        if (rc) return rc;
    }
    const int64_t prep_blocks = cdiv(A.ntiles, GZ_XCDS) * GZ_XCDS * cdiv(F, GZ_PREP_FG);
    if (prep_blocks > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_realize: too many tiles x frequencies");
    if ((int64_t)A.nchunks * F > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_realize: too many chunks x frequencies");
    hipLaunchKernelGGL(k_gz_prep, dim3((unsigned)prep_blocks), dim3(GZ_TILE), 0, st, A);
    int rc = check_launch("k_gz_prep");
    if (rc) return rc;
    hipLaunchKernelGGL(k_gz_plan, dim3(1), dim3(GZ_PLAN_THREADS), 0, st, A);
    rc = check_launch("k_gz_plan");
    if (rc) return rc;
    const dim3 grid(A.nchunks, F);
    const size_t lds = GZ_DRAW_LDS;
    if (mode == HD_SEL_SORTED) {
        if (L == 1) rc = gz_launch_np<HD_SEL_SORTED, 1>(A, npar, grid, bs, lds, st);
        else if (L == 5) rc = gz_launch_np<HD_SEL_SORTED, 5>(A, npar, grid, bs, lds, st);
        else if (L == 10) rc = gz_launch_np<HD_SEL_SORTED, 10>(A, npar, grid, bs, lds, st);
        else if (L <= 4) rc = gz_launch_np<HD_SEL_SORTED, 4>(A, npar, grid, bs, lds, st);
        else rc = gz_launch_np<HD_SEL_SORTED, 16>(A, npar, grid, bs, lds, st);
    } else if (mode == HD_SEL_ARGMAX) {
        if (L == 1) rc = gz_launch_np<HD_SEL_ARGMAX, 1>(A, npar, grid, bs, lds, st);
        else if (L <= 4) rc = gz_launch_np<HD_SEL_ARGMAX, 4>(A, npar, grid, bs, lds, st);
        else rc = gz_launch_np<HD_SEL_ARGMAX, 16>(A, npar, grid, bs, lds, st);
    } else {
        rc = gz_launch_np<HD_SEL_NONE, 1>(A, npar, grid, bs, lds, st);
    }
    if (rc) return rc;
    return realize_merge_launch(h2fdf, par, order, B, F, R, L, mode, npar, A.nchunks, A.p_total, A.p_ptot, A.p_key,
                                A.p_bin, hc2ss, ssidx, hc2bg, (npar > 0) ? bgpar : nullptr, group_ws, st);
}

extern "C" int hd_realize(const double* number, const double* h2fdf, const int64_t* order, int64_t B, int F, int R, int L,
                          int mode, int flags, double normal_threshold, uint64_t seed, const double* par, int npar,
                          double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar, void* workspace,
                          size_t ws_bytes, void* stream) {
    return realize_impl(number, h2fdf, order, B, F, R, L, mode, flags, normal_threshold, seed, par, npar, hc2ss, ssidx,
                        hc2bg, bgpar, nullptr, workspace, ws_bytes, stream);
}

extern "C" int hd_realize_record_counts(const double* number, const double* h2fdf, const int64_t* order, int64_t B,
                                        int F, int R, int L, int mode, int flags, double normal_threshold,
                                        uint64_t seed, const double* par, int npar, double* hc2ss, int64_t* ssidx,
                                        double* hc2bg, double* bgpar, double* counts_out, void* workspace,
                                        size_t ws_bytes, void* stream) {
    if (!counts_out) return set_error(HD_ERR_ARG, "hd_realize_record_counts: counts_out required");
    return realize_impl(number, h2fdf, order, B, F, R, L, mode, flags, normal_threshold, seed, par, npar, hc2ss, ssidx,
                        hc2bg, bgpar, counts_out, workspace, ws_bytes, stream);
}

// ---------------------------------------------------------------------------
// hd_realize_noalloc: hd_realize's launch sequence on a fixed device array
// ---------------------------------------------------------------------------
// The workspace of hd_realize (records, sparse axes, partials, merge stage) lives
// in this file-scope array instead of a caller allocation.  Every GPU the library
// runs on holds its own copy, HD_NOALLOC_SCRATCH_BYTES, until the process ends.
// Host code never passes `scratch` itself to a kernel (its host-side name is a
// stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) unsigned char scratch[HD_NOALLOC_SCRATCH_BYTES];

namespace hd {
constexpr int GZ_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[GZ_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up on the first call per device).
static int scratch_address(hipStream_t st, void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(HD_ERR_HIP, "hd_realize_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= GZ_MAX_DEVICES)
        return set_error(HD_ERR_ARG, "hd_realize_noalloc: device %d beyond %d", dev, GZ_MAX_DEVICES);
    if (st) {   // the scratch of the stream's device is the one the kernels will touch
        hipDevice_t sdev = -1;
        e = hipStreamGetDevice(st, &sdev);
        if (e != hipSuccess)
            return set_error(HD_ERR_HIP, "hd_realize_noalloc: hipStreamGetDevice: %s", hipGetErrorString(e));
        if (sdev != dev)
            return set_error(HD_ERR_ARG, "hd_realize_noalloc: stream of device %d, current device %d", (int)sdev, dev);
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(HD_ERR_HIP, "hd_realize_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return HD_OK;
}
}  // namespace hd

extern "C" int hd_realize_noalloc(const double* number, const double* h2fdf, const int64_t* order, int64_t B, int F,
                                  int R, int L, int mode, int flags, double normal_threshold, uint64_t seed,
                                  const double* par, int npar, double* hc2ss, int64_t* ssidx, double* hc2bg,
                                  double* bgpar, void* stream) {
    // the sizes are checked before the scratch is sized against them (gz_ws needs sane values)
    if (B < 1 || F < 1 || R < 1) return set_error(HD_ERR_ARG, "hd_realize_noalloc: bad arguments");
    if (B > 0x7fffffffll || F > 65535 || R > 0x7fffffff / 2)
        return set_error(HD_ERR_SIZE, "hd_realize_noalloc: sizes B=%lld F=%d R=%d out of range", (long long)B, F, R);
    const int Lw = (mode == HD_SEL_NONE) ? 0 : L;
    if (Lw < 0 || Lw > 16 || npar < 0 || npar > 7)
        return set_error(HD_ERR_ARG, "hd_realize_noalloc: L=%d npar=%d invalid", L, npar);
    const size_t need = gz_ws(B, F, R, Lw, npar);
    if (need > (size_t)HD_NOALLOC_SCRATCH_BYTES)
        return set_error(HD_ERR_SIZE,
                         "hd_realize_noalloc: B=%lld F=%d R=%d L=%d npar=%d needs %zu bytes of scratch, the fixed "
                         "array holds %zu; use hd_realize with a workspace",
                         (long long)B, F, R, L, npar, need, (size_t)HD_NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    int rc = scratch_address((hipStream_t)stream, &ws);
    if (rc) return rc;
    // Nothing in the scratch has to start at zero: every word the kernels read is written
    // earlier in the same launch sequence (prep: records, counts, sparse axes; draw: every
    // partial the merges read), so the previous call's contents are never seen.
    return realize_impl(number, h2fdf, order, B, F, R, L, mode, flags, normal_threshold, seed, par, npar, hc2ss, ssidx,
                        hc2bg, bgpar, nullptr, ws, (size_t)HD_NOALLOC_SCRATCH_BYTES, stream);
}

extern "C" int hd_poisson_window(double lam, int cap, int* kb, int* kt, int* n) {
    if (!(lam > 0.0) || !(lam < 1e300) || cap < 1 || !kb || !kt || !n)
        return set_error(HD_ERR_ARG, "hd_poisson_window: bad arguments");
    gz_window(lam, cap, *kb, *kt, *n);
    return HD_OK;
}

extern "C" int hd_poisson_cdf_table(double lam, int cap, int* kb_kt_n, uint64_t* cdf, void* stream) {
    if (!(lam > 0.0) || !(lam < 1e300) || cap < 1 || cap > GZ_HOOK_CAP_MAX || !kb_kt_n || !cdf)
        return set_error(HD_ERR_ARG, "hd_poisson_cdf_table: bad arguments");
    {   // the table is sized on the host first: a launch only when it fits
        int kb, kt, n;
        gz_window(lam, cap, kb, kt, n);
        if (n < 1) return set_error(HD_ERR_SIZE, "hd_poisson_cdf_table: table of lam=%g exceeds cap=%d", lam, cap);
    }
    hipLaunchKernelGGL(k_cdf_table_hook, dim3(1), dim3(64), (size_t)cap * sizeof(double), (hipStream_t)stream, lam,
                       cap, kb_kt_n, cdf);
    return check_launch("k_cdf_table_hook");
}

static bool gz_hook_table_ok(int n, double lam, int kb, int kt) {
    if (n < 2 || n > GZ_HOOK_CAP_MAX || !(lam > 0.0) || kb < 0 || kt < kb) return false;
    int b, t, w;   // the table must be the one the kernels build for lam (its tails walk from there)
    gz_window(lam, GZ_HOOK_CAP_MAX, b, t, w);
    return w == n && b == kb && t == kt;
}

extern "C" int hd_table_draw_words(const uint32_t* tab, int n, double lam, int kb, int kt, const uint32_t* words,
                                   int64_t m, int* count, int* used, void* stream) {
    if (!tab || !gz_hook_table_ok(n, lam, kb, kt) || m < 0 || (m > 0 && (!words || !count || !used)))
        return set_error(HD_ERR_ARG, "hd_table_draw_words: bad arguments");
    if (m == 0) return HD_OK;
    hipLaunchKernelGGL(k_table_draw_hook, dim3((unsigned)cdiv(m, 256)), dim3(256), (size_t)n * 2 * sizeof(uint32_t),
                       (hipStream_t)stream, tab, n, lam, kb, kt, words, m, count, used);
    return check_launch("k_table_draw_hook");
}

extern "C" void hd_xoshiro128ss(uint32_t* state, int64_t n, uint32_t* out) {
    Xo128 g;
    g.s0 = state[0];
    g.s1 = state[1];
    g.s2 = state[2];
    g.s3 = state[3];
    for (int64_t i = 0; i < n; i++) {
        const uint32_t v = g.next();
        if (out) out[i] = v;
    }
    state[0] = g.s0;
    state[1] = g.s1;
    state[2] = g.s2;
    state[3] = g.s3;
}

extern "C" int hd_table_draw_pair_words(const uint32_t* tab, int n, double lam, int kb, int kt, const uint32_t* tab2,
                                        int n2, double lam2, int kb2, int kt2, const uint32_t* words, int64_t m,
                                        int* count, int* used, void* stream) {
    if (!tab || !tab2 || !gz_hook_table_ok(n, lam, kb, kt) || !gz_hook_table_ok(n2, lam2, kb2, kt2) ||
        n + n2 > GZ_HOOK_CAP_MAX || m < 0 || (m > 0 && (!words || !count || !used)))
        return set_error(HD_ERR_ARG, "hd_table_draw_pair_words: bad arguments");
    if (m == 0) return HD_OK;
    hipLaunchKernelGGL(k_table_draw_pair_hook, dim3((unsigned)cdiv(m, 256)), dim3(256),
                       (size_t)(n + n2) * 2 * sizeof(uint32_t), (hipStream_t)stream, tab, n, lam, kb, kt, tab2, n2,
                       lam2, kb2, kt2, words, m, count, used);
    return check_launch("k_table_draw_pair_hook");
}
