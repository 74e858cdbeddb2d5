// mlp.hip -- kernels of the learned GWB surrogate: a small MLP (parameters ->
// log10 h_c percentiles) trained and evaluated on the GPU.
//
// No counterpart in the reference (its emulator is a Gaussian process,
// holodeck/gps/); this is the surrogate that sits beside it.  Arithmetic:
// BF16 operands on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), FP32
// accumulation, FP32 master weights and Adam moments.
//
// GEMM: C (M x N) = A (M x K) . B^T with B given as (N x K); both operands are
// read through (row, k) strides, so one kernel serves the forward pass
// (activations . W^T), the input gradient (dZ . W) and the weight gradient
// (dZ^T . activations).  64 x 64 output tile per 256-thread workgroup, 4 waves
// of 32 x 32 (2 x 2 MFMA tiles of 16 x 16), K in steps of 32 staged in LDS.
//
// MFMA 16x16x32 operand layout (lane l of 64): A row l%16, k 8*(l/16)..+7;
// B column l%16, k 8*(l/16)..+7; accumulator column l%16, rows 4*(l/16)..+3.
#include "common.hpp"
#include "../../include/holohip.h"

namespace hd {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int MT = 64;       // tile M
constexpr int NT = 64;       // tile N
constexpr int KT = 32;       // tile K
constexpr int KP = KT + 8;   // padded LDS row (bf16): 80 bytes, keeps 16-byte alignment

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Epilogues
//   HD_MLP_EPI_STORE      out_f32[m, n] = acc (+ bias[n])
//   HD_MLP_EPI_BIAS_SILU  z = acc + bias[n]; out_f32[m, n] = z; out_bf16[m, n] = silu(z)
//   HD_MLP_EPI_BIAS_BF16  out_bf16[m, n] = acc + bias[n]  (and out_f32 if given)
//   HD_MLP_EPI_SILU_BWD   out_bf16[m, n] = acc * silu'(zin[m, n])
template <int EPI>
__global__ void __launch_bounds__(256) k_mlp_gemm(const bf16* __restrict__ A, int64_t sam, int64_t sak,
                                                  const bf16* __restrict__ Bm, int64_t sbn, int64_t sbk, int M, int N,
                                                  int K, const float* __restrict__ bias,
                                                  const float* __restrict__ zin, float* __restrict__ out_f32,
                                                  bf16* __restrict__ out_bf16, int64_t ldo) {
    __shared__ __attribute__((aligned(16))) bf16 As[MT][KP];
    __shared__ __attribute__((aligned(16))) bf16 Bs[NT][KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * MT, n0 = blockIdx.x * NT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // staging: thread tid loads 8 consecutive k of one row of each tile
    const int lr = tid >> 2;          // 0..63
    const int lk = (tid & 3) * 8;     // 0, 8, 16, 24
    const bool a_vec = (sak == 1) && ((sam & 7) == 0) && ((((uintptr_t)A) & 15) == 0);
    const bool b_vec = (sbk == 1) && ((sbn & 7) == 0) && ((((uintptr_t)Bm) & 15) == 0);
    for (int k0 = 0; k0 < K; k0 += KT) {
        {
            const int gm = m0 + lr;
            bf16x8 v;
            if (a_vec && gm < M && k0 + lk + 8 <= K) {
                v = *(const bf16x8*)(A + gm * sam + (k0 + lk));
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int gk = k0 + lk + j;
                    v[j] = (gm < M && gk < K) ? A[gm * sam + gk * sak] : (bf16)0.0f;
                }
            }
            *(bf16x8*)&As[lr][lk] = v;
        }
        {
            const int gn = n0 + lr;
            bf16x8 v;
            if (b_vec && gn < N && k0 + lk + 8 <= K) {
                v = *(const bf16x8*)(Bm + gn * sbn + (k0 + lk));
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int gk = k0 + lk + j;
                    v[j] = (gn < N && gk < K) ? Bm[gn * sbn + gk * sbk] : (bf16)0.0f;
                }
            }
            *(bf16x8*)&Bs[lr][lk] = v;
        }
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&As[wm + 16 * i + fr][fk];
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&Bs[wn + 16 * j + fr][fk];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    const int cn = lane & 15, cm = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wm + 16 * i + cm + r;
                const int gn = n0 + wn + 16 * j + cn;
                if (gm >= M || gn >= N) continue;
                float x = acc[i][j][r];
                const int64_t o = (int64_t)gm * ldo + gn;
                if (EPI == HD_MLP_EPI_STORE) {
                    if (bias) x += bias[gn];
                    out_f32[o] = x;
                } else if (EPI == HD_MLP_EPI_BIAS_SILU) {
                    x += bias[gn];
                    if (out_f32) out_f32[o] = x;
                    out_bf16[o] = (bf16)silu(x);
                } else if (EPI == HD_MLP_EPI_BIAS_BF16) {
                    if (bias) x += bias[gn];
                    if (out_f32) out_f32[o] = x;
                    out_bf16[o] = (bf16)x;
                } else {   // HD_MLP_EPI_SILU_BWD
                    out_bf16[o] = (bf16)(x * silu_grad(zin[o]));
                }
            }
}

// column sums of a bf16 (M x N) matrix into FP32 (bias gradients); one thread per column
__global__ void __launch_bounds__(256) k_mlp_colsum(const bf16* __restrict__ a, int M, int N, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; m++) s += (float)a[(int64_t)m * N + n];
    out[n] = s;
}

// mean-squared error over (M x N): dout = 2 (y - t) / (M N) in bf16; loss accumulated (FP32)
__global__ void __launch_bounds__(256) k_mlp_mse(const float* __restrict__ y, const float* __restrict__ t, int64_t n,
                                                 float scale, bf16* __restrict__ dout, float* __restrict__ loss) {
    __shared__ float s_part[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (i < n) {
        const float d = y[i] - t[i];
        dout[i] = (bf16)(2.0f * d * scale);
        e = d * d;
    }
    s_part[threadIdx.x] = e;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s_part[threadIdx.x] += s_part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, s_part[0] * scale);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the bf16 copy
__global__ void __launch_bounds__(256) k_mlp_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2, bf16* __restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = (bf16)wi;
}

__global__ void __launch_bounds__(256) k_f32_to_bf16(const float* __restrict__ a, int64_t n, bf16* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (bf16)a[i];
}

template <int EPI>
static void gemm_launch(dim3 grid, hipStream_t st, const bf16* A, int64_t sam, int64_t sak, const bf16* B, int64_t sbn,
                        int64_t sbk, int M, int N, int K, const float* bias, const float* zin, float* of, bf16* ob,
                        int64_t ldo) {
    hipLaunchKernelGGL((k_mlp_gemm<EPI>), grid, dim3(256), 0, st, A, sam, sak, B, sbn, sbk, M, N, K, bias, zin, of, ob,
                       ldo);
}

}  // namespace hd

using namespace hd;

extern "C" int hd_mlp_gemm_bf16(const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbn, int64_t sbk, int M,
                                int N, int K, int epi, const float* bias, const float* zin, float* out_f32,
                                void* out_bf16, int64_t ldo, void* stream) {
    if (!A || !B || M < 1 || N < 1 || K < 1 || ldo < N) return set_error(HD_ERR_ARG, "hd_mlp_gemm_bf16: bad arguments");
    if (cdiv(M, MT) > 65535) return set_error(HD_ERR_SIZE, "hd_mlp_gemm_bf16: M=%d too large (> %d)", M, 65535 * MT);
    const bf16* a = (const bf16*)A;
    const bf16* b = (const bf16*)B;
    bf16* ob = (bf16*)out_bf16;
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)cdiv(N, NT), (unsigned)cdiv(M, MT));
    KTimer timer("k_mlp_gemm", st);
    switch (epi) {
        case HD_MLP_EPI_STORE:
            if (!out_f32) return set_error(HD_ERR_ARG, "hd_mlp_gemm_bf16: STORE needs out_f32");
            gemm_launch<HD_MLP_EPI_STORE>(grid, st, a, sam, sak, b, sbn, sbk, M, N, K, bias, zin, out_f32, ob, ldo);
            break;
        case HD_MLP_EPI_BIAS_SILU:
            if (!bias || !ob) return set_error(HD_ERR_ARG, "hd_mlp_gemm_bf16: BIAS_SILU needs bias and out_bf16");
            gemm_launch<HD_MLP_EPI_BIAS_SILU>(grid, st, a, sam, sak, b, sbn, sbk, M, N, K, bias, zin, out_f32, ob, ldo);
            break;
        case HD_MLP_EPI_BIAS_BF16:
            if (!ob) return set_error(HD_ERR_ARG, "hd_mlp_gemm_bf16: BIAS_BF16 needs out_bf16");
            gemm_launch<HD_MLP_EPI_BIAS_BF16>(grid, st, a, sam, sak, b, sbn, sbk, M, N, K, bias, zin, out_f32, ob, ldo);
            break;
        case HD_MLP_EPI_SILU_BWD:
            if (!zin || !ob) return set_error(HD_ERR_ARG, "hd_mlp_gemm_bf16: SILU_BWD needs zin and out_bf16");
            gemm_launch<HD_MLP_EPI_SILU_BWD>(grid, st, a, sam, sak, b, sbn, sbk, M, N, K, bias, zin, out_f32, ob, ldo);
            break;
        default:
            return set_error(HD_ERR_ARG, "hd_mlp_gemm_bf16: unknown epilogue %d", epi);
    }
    return check_launch("k_mlp_gemm");
}

extern "C" int hd_mlp_colsum_bf16(const void* a, int M, int N, float* out, void* stream) {
    if (!a || !out || M < 1 || N < 1) return set_error(HD_ERR_ARG, "hd_mlp_colsum_bf16: bad arguments");
    hipLaunchKernelGGL(k_mlp_colsum, dim3((unsigned)cdiv(N, 256)), dim3(256), 0, (hipStream_t)stream, (const bf16*)a, M,
                       N, out);
    return check_launch("k_mlp_colsum");
}

extern "C" int hd_mlp_mse(const float* y, const float* t, int64_t n, void* dout, float* loss, void* stream) {
    if (!y || !t || !dout || !loss || n < 1) return set_error(HD_ERR_ARG, "hd_mlp_mse: bad arguments");
    hipLaunchKernelGGL(k_mlp_mse, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, y, t, n,
                       1.0f / (float)n, (bf16*)dout, loss);
    return check_launch("k_mlp_mse");
}

extern "C" int hd_mlp_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                           float eps, int step, void* w_bf16, void* stream) {
    if (!w || !g || !m || !v || n < 1 || step < 1) return set_error(HD_ERR_ARG, "hd_mlp_adam: bad arguments");
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, w, g, m, v, lr, b1,
                       b2, eps, c1, c2, (bf16*)w_bf16);
    return check_launch("k_mlp_adam");
}

extern "C" int hd_f32_to_bf16(const float* a, int64_t n, void* b, void* stream) {
    if (!a || !b || n < 1) return set_error(HD_ERR_ARG, "hd_f32_to_bf16: bad arguments");
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, a, n, (bf16*)b);
    return check_launch("k_f32_to_bf16");
}
