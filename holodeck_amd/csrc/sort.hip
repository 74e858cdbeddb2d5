// sort.hip -- stable descending order of bins by an FP64 key.
//
// <- `indices = np.argsort(-h2fdf[..., 0].flatten())` (single_sources.py:89).
// numpy's default argsort is not stable; this order is: key descending, ties by
// ascending bin index (which is also what a stable sort of -key gives).
//
// LSD radix sort over the 64-bit order-preserving image of the key, 8 bits per
// pass.  Each pass: per-tile digit histogram, one exclusive scan of the
// (digit-major, tile-minor) counts, and a stable scatter in which each of a
// tile's four wavefronts walks a quarter of it in chunks of 64 and ranks equal
// digits with ballots.  Passes whose digit is the same for every key are skipped on the
// device (the scatter kernel copies).
#include "common.hpp"
#include "../../include/holohip.h"

namespace hd {

constexpr int SORT_TILE = 4096;     // keys per tile
constexpr int SORT_WAVES = 4;       // waves per tile: each ranks a contiguous quarter
constexpr int SORT_RADIX = 256;
constexpr int SORT_SCAN_THREADS = 1024;

__device__ __forceinline__ uint64_t desc_key(double v) {
    if (v == 0.0) v = 0.0;   // -0 -> +0
    uint64_t u = (uint64_t)__double_as_longlong(v);
    u = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);   // ascending image
    return ~u;   // descending
}

__global__ void k_sort_init(const double* __restrict__ key, int64_t B, int64_t stride, int64_t col,
                            uint64_t* __restrict__ k_out, int64_t* __restrict__ v_out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    k_out[i] = desc_key(key[i * stride + col]);
    v_out[i] = i;
}

// digit histogram of one tile (256 threads, LDS atomics), stored digit-major
__global__ void __launch_bounds__(256) k_sort_hist(const uint64_t* __restrict__ keys, int64_t B, int shift,
                                                   uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t h[SORT_RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    const int64_t end = min(base + SORT_TILE, B);
    for (int64_t i = base + threadIdx.x; i < end; i += 256) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xff], 1u);
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of n counts (16-byte aligned) by one workgroup: each thread sums a contiguous
// segment (four counts per load), the segment sums are scanned in LDS, and each thread rewrites
// its segment; also flags whether a single digit holds every key
__global__ void __launch_bounds__(SORT_SCAN_THREADS) k_sort_scan(uint32_t* __restrict__ data, int64_t n, int ntiles,
                                                                 int64_t B, int* __restrict__ trivial) {
    __shared__ uint32_t s_part[SORT_SCAN_THREADS];
    __shared__ int s_triv;
    const int tid = threadIdx.x;
    const int64_t per = ((n + SORT_SCAN_THREADS - 1) / SORT_SCAN_THREADS + 3) & ~(int64_t)3;
    const int64_t lo = min((int64_t)tid * per, n), hi = min(lo + per, n);
    const int64_t hi4 = lo + ((hi - lo) & ~(int64_t)3);
    uint32_t acc = 0;
#pragma unroll 4
    for (int64_t i = lo; i < hi4; i += 4) {
        const uint4 v = *(const uint4*)(data + i);
        acc += v.x + v.y + v.z + v.w;
    }
    for (int64_t i = hi4; i < hi; i++) acc += data[i];
    s_part[tid] = acc;
    if (tid == 0) s_triv = 0;
    __syncthreads();
    for (int off = 1; off < SORT_SCAN_THREADS; off <<= 1) {   // Hillis-Steele over the segment sums
        const uint32_t v = (tid >= off) ? s_part[tid - off] : 0u;
        __syncthreads();
        s_part[tid] += v;
        __syncthreads();
    }
    uint32_t run = (tid == 0) ? 0u : s_part[tid - 1];
#pragma unroll 4
    for (int64_t i = lo; i < hi4; i += 4) {
        uint4 v = *(uint4*)(data + i);
        uint4 o;
        o.x = run;
        run += v.x;
        o.y = run;
        run += v.y;
        o.z = run;
        run += v.z;
        o.w = run;
        run += v.w;
        *(uint4*)(data + i) = o;
    }
    for (int64_t i = hi4; i < hi; i++) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
    __syncthreads();
    // trivial pass: some digit d has offsets [d*ntiles] == 0 and [(d+1)*ntiles] == B
    for (int d = tid; d < SORT_RADIX; d += SORT_SCAN_THREADS) {
        const uint32_t start = data[(int64_t)d * ntiles];
        const uint32_t stop = (d + 1 < SORT_RADIX) ? data[(int64_t)(d + 1) * ntiles] : (uint32_t)B;
        if (start == 0 && stop == (uint32_t)B) s_triv = 1;
    }
    __syncthreads();
    if (tid == 0) *trivial = s_triv;
}

// Stable scatter of one tile by four waves, each ranking a contiguous quarter in chunks of 64
// keys with ballots; a wave's digit offsets are the tile's plus the counts of the quarters
// before it.  Within a wave the running offsets are read and then advanced in program order
// (LDS operations of a wave complete in order), so no workgroup barrier is taken per chunk.
__global__ void __launch_bounds__(64 * SORT_WAVES) k_sort_scatter(const uint64_t* __restrict__ kin,
                                                                  const int64_t* __restrict__ vin,
                                                                  uint64_t* __restrict__ kout,
                                                                  int64_t* __restrict__ vout, int64_t B, int shift,
                                                                  const uint32_t* __restrict__ offs, int ntiles,
                                                                  const int* __restrict__ trivial) {
    __shared__ uint32_t s_h[SORT_WAVES][SORT_RADIX];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    const int64_t end = min(base + SORT_TILE, B);
    if (*trivial) {
        for (int64_t i = base + tid; i < end; i += 64 * SORT_WAVES) {
            kout[i] = kin[i];
            vout[i] = vin[i];
        }
        return;
    }
    constexpr int QUARTER = SORT_TILE / SORT_WAVES;
    const int64_t wb = base + (int64_t)w * QUARTER, we = min(wb + QUARTER, end);
    for (int d = lane; d < SORT_RADIX; d += 64) s_h[w][d] = 0;
    __syncthreads();
    for (int64_t i = wb + lane; i < we; i += 64) atomicAdd(&s_h[w][(uint32_t)(kin[i] >> shift) & 0xff], 1u);
    __syncthreads();
    {   // thread d: the offsets of digit d for the four quarters
        const int d = tid;
        uint32_t run = offs[(int64_t)d * ntiles + blockIdx.x];
#pragma unroll
        for (int q = 0; q < SORT_WAVES; q++) {
            const uint32_t c = s_h[q][d];
            s_h[q][d] = run;
            run += c;
        }
    }
    __syncthreads();
    uint32_t* run = s_h[w];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int64_t c = wb; c < we; c += 64) {
        const int64_t i = c + lane;
        const bool valid = i < we;
        const uint64_t k = valid ? kin[i] : 0;
        const uint32_t d = (uint32_t)(k >> shift) & 0xff;
        uint64_t eq = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t b = __ballot(valid && ((d >> bit) & 1));
            eq &= ((d >> bit) & 1) ? b : ~b;
        }
        const uint32_t rank = (uint32_t)__popcll(eq & lt_mask);
        const uint32_t cnt = (uint32_t)__popcll(eq);
        uint32_t pos = 0;
        if (valid) pos = run[d] + rank;
        asm volatile("" ::: "memory");   // every lane's read of run[] before any advance
        if (valid && rank == cnt - 1) run[d] = pos + 1;
        asm volatile("" ::: "memory");
        if (valid) {
            kout[pos] = k;
            vout[pos] = vin[i];
        }
    }
}

}  // namespace hd

using namespace hd;

extern "C" size_t hd_order_workspace_size(int64_t B) {
    int64_t ntiles = cdiv(B, SORT_TILE);
    size_t s = 0;
    s += 2 * (size_t)B * sizeof(uint64_t);   // keys ping-pong (2nd buffer)
    s += (size_t)B * sizeof(int64_t);        // values ping buffer
    s += (size_t)SORT_RADIX * ntiles * sizeof(uint32_t) + 256;   // counts (256-byte aligned)
    s += 256;                                // trivial flag + alignment
    return s + 1024;
}

extern "C" int hd_order_desc(const double* key, int64_t B, int64_t stride, int64_t col, int64_t* order, void* workspace,
                             size_t ws_bytes, void* stream) {
    if (B < 1 || stride < 1 || col < 0 || col >= stride || !key || !order)
        return set_error(HD_ERR_ARG, "hd_order_desc: bad arguments");
    if (B > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_order_desc: B too large");
    if (!workspace || ws_bytes < hd_order_workspace_size(B))
        return set_error(HD_ERR_SIZE, "hd_order_desc: workspace too small (%zu < %zu)", ws_bytes,
                         hd_order_workspace_size(B));
    hipStream_t st = (hipStream_t)stream;
    int ntiles = (int)cdiv(B, SORT_TILE);
    char* p = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    uint64_t* k0 = (uint64_t*)p;            p += (size_t)B * 8;
    uint64_t* k1 = (uint64_t*)p;            p += (size_t)B * 8;
    int64_t* v1 = (int64_t*)p;              p += (size_t)B * 8;
    p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    uint32_t* hist = (uint32_t*)p;          p += (size_t)SORT_RADIX * ntiles * 4;
    int* triv = (int*)(((uintptr_t)p + 15) & ~(uintptr_t)15);
    int64_t* v0 = order;   // values ping-pong through the output buffer; 8 passes end in v0

    hipLaunchKernelGGL(k_sort_init, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, st, key, B, stride, col, k0, v0);
    int rc = check_launch("k_sort_init");
    if (rc) return rc;
    uint64_t* kin = k0; uint64_t* kout = k1;
    int64_t* vin = v0; int64_t* vout = v1;
    for (int pass = 0; pass < 8; pass++) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(256), 0, st, kin, B, shift, hist, ntiles);
        hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(SORT_SCAN_THREADS), 0, st, hist, (int64_t)SORT_RADIX * ntiles,
                           ntiles, B, triv);
        hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(64 * SORT_WAVES), 0, st, kin, vin, kout, vout, B, shift,
                           hist, ntiles, triv);
        rc = check_launch("k_sort pass");
        if (rc) return rc;
        uint64_t* tk = kin; kin = kout; kout = tk;
        int64_t* tv = vin; vin = vout; vout = tv;
    }
    // after an even number of passes the result is back in k0 / v0 (= order)
    return HD_OK;
}
