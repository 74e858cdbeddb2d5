// scatter.hip -- M-Mbulge scatter of the binary density (FP64), all redshift slices at once.
//
// <- add_scatter_to_masses (holodeck/sams/sam.py:1291-1394).  Per slice:
//   1. global gradient estimation of the Clough-Tocher interpolant at the
//      (log10 m1, log10 m2) nodes: the Gauss-Seidel minimisation of the edge
//      curvature (G. M. Nielson, Math. Comp. 40 (1983) 253; R. J. Renka and
//      A. K. Cline, Rocky Mountain J. Math. 14 (1984) 223), tol 1e-6, at most
//      400 sweeps, as scipy's CloughTocher2DInterpolator defaults;
//   2. Clough-Tocher evaluation on the regular (G x G) log-mass grid (cubic
//      Bernstein-Bezier patches on the three micro-triangles, affine-invariant
//      cross-boundary derivatives; P. Alfeld, CAGD 1 (1984) 169; G. Farin,
//      CAGD 3 (1986) 83), NaN/negative -> nearest node value;
//   3. Y = P D with P = (W W)^T, the reference's two einsum passes (both act on
//      the m1 axis), as one FP64 GEMM over all slices;
//   4. bilinear interpolation back to the nodes.
//
// The Gauss-Seidel sweep is level-scheduled: a vertex's level is one more than
// the highest level of its lower-indexed neighbours, so updating whole levels
// in order reads exactly the values the sequential sweep reads (new values for
// lower indices, previous-sweep values for higher ones).  One wavefront runs
// one slice with the solution vector in LDS.
#include "common.hpp"
#include "../../include/holohip.h"

#pragma clang fp contract(off)

namespace hd {

// Slot order: slot t is position t of the level schedule (vertex lidx[t]).
// slot_eoff: exclusive prefix of the slots' degrees (edges of slot t are
// [slot_eoff[t], slot_eoff[t+1]) in slot-ordered CSR).  One workgroup.
__global__ void __launch_bounds__(1024) k_ct_slots(int N, const int32_t* __restrict__ nptr,
                                                   const int32_t* __restrict__ lidx, int32_t* __restrict__ slot_eoff) {
    __shared__ int32_t s_part[1024];
    const int tid = threadIdx.x;
    int32_t carry = 0;
    for (int base = 0; base < N; base += 1024) {
        const int t = base + tid;
        const int32_t d = (t < N) ? nptr[lidx[t] + 1] - nptr[lidx[t]] : 0;
        s_part[tid] = d;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int32_t v = (tid >= off) ? s_part[tid - off] : 0;
            __syncthreads();
            s_part[tid] += v;
            __syncthreads();
        }
        if (t < N) slot_eoff[t] = carry + s_part[tid] - d;
        carry += s_part[1023];
        __syncthreads();
    }
    if (tid == 0) slot_eoff[N] = carry;
}

// Geometry of every slot: the 2x2 system (Q0, Q1, Q3, det) and the edges
// (neighbour, ex, ey, L^3) in slot-ordered CSR -- the same expressions in the
// same order as the sweep used to evaluate them, so the sweep's arithmetic is
// unchanged.  Edge storage holds ecap entries (6 N bounds a planar graph's
// directed edges); an overflow is reported through *bad.
__global__ void __launch_bounds__(256) k_ct_geom(int N, const double* __restrict__ pts, const int32_t* __restrict__ nptr,
                                                 const int32_t* __restrict__ nidx, const int32_t* __restrict__ lidx,
                                                 const int32_t* __restrict__ slot_eoff, int64_t ecap,
                                                 double* __restrict__ slot_q, int32_t* __restrict__ e_q,
                                                 double* __restrict__ e_geo, int32_t* __restrict__ bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int p = lidx[t];
    const double px = pts[2 * p], py = pts[2 * p + 1];
    double Q0 = 0, Q1 = 0, Q3 = 0;
    int64_t e = slot_eoff[t];
    if ((int64_t)slot_eoff[N] > ecap) {
        if (t == 0) atomicAdd(bad, 1 << 20);
        return;
    }
    for (int jj = nptr[p]; jj < nptr[p + 1]; jj++, e++) {
        const int q = nidx[jj];
        const double ex = pts[2 * q] - px;
        const double ey = pts[2 * q + 1] - py;
        const double L = sqrt(ex * ex + ey * ey);
        const double L3 = L * L * L;
        Q0 += 4 * ex * ex / L3;
        Q1 += 4 * ex * ey / L3;
        Q3 += 4 * ey * ey / L3;
        e_q[e] = q;
        e_geo[3 * e] = ex;
        e_geo[3 * e + 1] = ey;
        e_geo[3 * e + 2] = L3;
    }
    const double Q2 = Q1;
    slot_q[4 * t] = Q0;
    slot_q[4 * t + 1] = Q1;
    slot_q[4 * t + 2] = Q3;
    slot_q[4 * t + 3] = Q0 * Q3 - Q1 * Q2;
}

// dens (N, Z) -> densT (Z, N): each slice contiguous
__global__ void __launch_bounds__(256) k_transpose_nz(const double* __restrict__ a, int N, int Z, double* __restrict__ b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * Z) return;
    const int64_t p = t / Z, z = t % Z;
    b[z * N + p] = a[t];
}

// Fixed-width copy of the slot-ordered geometry: the first CT_D edges of every
// slot in [d][slot] arrays (padding: neighbour 0, zero edge, L^3 = 1, never read
// as the slot's degree stops the loop), so a sweep level loads each slot's edges
// with independent, coalesced reads instead of chasing slot_eoff.
constexpr int CT_D = 8;
__global__ void __launch_bounds__(256) k_ct_pad(int N, const int32_t* __restrict__ slot_eoff,
                                                const int32_t* __restrict__ e_q, const double* __restrict__ e_geo,
                                                int32_t* __restrict__ s_deg, int32_t* __restrict__ pq,
                                                double* __restrict__ pex, double* __restrict__ pey,
                                                double* __restrict__ pl3) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int e0 = slot_eoff[t], deg = slot_eoff[t + 1] - e0;
    s_deg[t] = deg;
#pragma unroll
    for (int d = 0; d < CT_D; d++) {
        const bool in = d < deg;
        pq[d * N + t] = in ? e_q[e0 + d] : 0;
        pex[d * N + t] = in ? e_geo[3 * (e0 + d)] : 0.0;
        pey[d * N + t] = in ? e_geo[3 * (e0 + d) + 1] : 0.0;
        pl3[d * N + t] = in ? e_geo[3 * (e0 + d) + 2] : 1.0;
    }
}

// The per-slice constant of every padded edge, 6 (f1 - f2) (the sweep's first
// term, the same operation on the same values), A[z][d][slot].
__global__ void __launch_bounds__(256) k_ct_fdiff(const double* __restrict__ densT, int N, int Z,
                                                  const int32_t* __restrict__ lidx, const int32_t* __restrict__ s_deg,
                                                  const int32_t* __restrict__ pq, double* __restrict__ A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)N * Z) return;
    const int z = (int)(i / N), t = (int)(i % N);
    const double* frow = densT + (int64_t)z * N;
    const double f1 = frow[lidx[t]];
    const int deg = s_deg[t];
#pragma unroll
    for (int d = 0; d < CT_D; d++)
        A[((int64_t)z * CT_D + d) * N + t] = (d < deg) ? 6 * (f1 - frow[pq[d * N + t]]) : 0.0;
}

// One slot's inputs to a sweep step, held in registers (loaded one level ahead).
struct CtSlot {
    int p, deg, e0;
    double Q0, Q1, Q3, det, f1;
    int q[CT_D];
    double ex[CT_D], ey[CT_D], l3[CT_D], a[CT_D];
};

__device__ __forceinline__ void ct_load(CtSlot& s, int t, bool valid, int N, const double* __restrict__ frow,
                                        const double* __restrict__ Az, const int32_t* __restrict__ lidx,
                                        const int32_t* __restrict__ s_deg, const int32_t* __restrict__ slot_eoff,
                                        const double* __restrict__ slot_q, const int32_t* __restrict__ pq,
                                        const double* __restrict__ pex, const double* __restrict__ pey,
                                        const double* __restrict__ pl3) {
    s.deg = 0;
    if (!valid) return;
    s.p = lidx[t];
    s.deg = s_deg[t];
    s.e0 = slot_eoff[t];
    s.Q0 = slot_q[4 * t];
    s.Q1 = slot_q[4 * t + 1];
    s.Q3 = slot_q[4 * t + 2];
    s.det = slot_q[4 * t + 3];
    s.f1 = frow[s.p];
#pragma unroll
    for (int d = 0; d < CT_D; d++) {
        s.q[d] = pq[d * N + t];
        s.ex[d] = pex[d * N + t];
        s.ey[d] = pey[d * N + t];
        s.l3[d] = pl3[d * N + t];
        s.a[d] = Az[(int64_t)d * N + t];
    }
}

// One Gauss-Seidel update of slot s: scipy's arithmetic, term by term and in the
// same order (the first CT_D edges from registers, the rest of a wider slot from
// the slot-ordered CSR).  Returns the relative change.
__device__ __forceinline__ double ct_step(const CtSlot& c, double* y, const double* __restrict__ frow,
                                          const int32_t* __restrict__ e_q, const double* __restrict__ e_geo) {
    double s0 = 0, s1 = 0;
#pragma unroll
    for (int d = 0; d < CT_D; d++) {
        if (d < c.deg) {
            const int q = c.q[d];
            const double ex = c.ex[d], ey = c.ey[d], L3 = c.l3[d];
            const double df2 = -ex * y[2 * q] - ey * y[2 * q + 1];
            s0 += (c.a[d] - 2 * df2) * ex / L3;
            s1 += (c.a[d] - 2 * df2) * ey / L3;
        }
    }
    for (int e = c.e0 + CT_D; e < c.e0 + c.deg; e++) {
        const int q = e_q[e];
        const double ex = e_geo[3 * e], ey = e_geo[3 * e + 1], L3 = e_geo[3 * e + 2];
        const double f2 = frow[q];
        const double df2 = -ex * y[2 * q] - ey * y[2 * q + 1];
        s0 += (6 * (c.f1 - f2) - 2 * df2) * ex / L3;
        s1 += (6 * (c.f1 - f2) - 2 * df2) * ey / L3;
    }
    const double Q0 = c.Q0, Q1 = c.Q1, Q3 = c.Q3, det = c.det;
    const double Q2 = Q1;
    const double r0 = (Q3 * s0 - Q1 * s1) / det;
    const double r1 = (-Q2 * s0 + Q0 * s1) / det;
    const int p = c.p;
    double a0 = fabs(y[2 * p] + r0), a1 = fabs(y[2 * p + 1] + r1);
    double change = (a0 > a1) ? a0 : a1;
    y[2 * p] = -r0;
    y[2 * p + 1] = -r1;
    double m0 = fabs(r0), m1 = fabs(r1);
    double mr = (m0 > m1) ? m0 : m1;
    change /= (1.0 > mr) ? 1.0 : mr;
    return change;
}

// Gauss-Seidel sweeps of one slice (one wave).  Each level's slot inputs are
// loaded while the previous level is computed (they do not depend on the
// solution), into the other of two register buffers, so a level costs its LDS
// reads of the neighbours' gradients and its arithmetic.  With the solution in
// LDS the single wave's LDS operations complete in issue order, so a level's
// reads see the previous level's writes without a barrier.
template <bool LDS_Y>
__global__ void __launch_bounds__(64) k_ct_gradients(const double* __restrict__ densT, int N, int Z,
                                                     const int32_t* __restrict__ lidx,
                                                     const int32_t* __restrict__ slot_eoff,
                                                     const double* __restrict__ slot_q, const int32_t* __restrict__ e_q,
                                                     const double* __restrict__ e_geo, const int32_t* __restrict__ s_deg,
                                                     const int32_t* __restrict__ pq, const double* __restrict__ pex,
                                                     const double* __restrict__ pey, const double* __restrict__ pl3,
                                                     const double* __restrict__ A, const int32_t* __restrict__ lptr,
                                                     int n_levels, double tol, int maxiter, double* __restrict__ grad_g,
                                                     int32_t* __restrict__ iters_out) {
    extern __shared__ double s_y[];
    const int z = blockIdx.x;
    const int lane = threadIdx.x;
    const double* frow = densT + (int64_t)z * N;
    const double* Az = A + (int64_t)z * CT_D * N;
    double* y = LDS_Y ? s_y : (grad_g + (int64_t)z * N * 2);
    // the level pointers next to the solution (or alone) in LDS
    int32_t* s_lptr = (int32_t*)(LDS_Y ? (s_y + 2 * (size_t)N) : s_y);
    for (int i = lane; i < 2 * N; i += 64) y[i] = 0.0;
    for (int i = lane; i <= n_levels; i += 64) s_lptr[i] = lptr[i];
    __syncthreads();
    auto load = [&](CtSlot& b, int lev) {
        const int t = s_lptr[lev] + lane;
        ct_load(b, t, t < s_lptr[lev + 1], N, frow, Az, lidx, s_deg, slot_eoff, slot_q, pq, pex, pey, pl3);
    };
    auto level_done = [&]() {
        if (LDS_Y) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __threadfence_block();
        }
    };
    CtSlot b0, b1;
    load(b0, 0);
    int iters = 0, it = 0, lev = 0;
    double err = 0.0;
    // the steps (sweep, level) in order, two per trip: b0 holds the even steps, b1 the odd
    // ones; returns true when a sweep ends converged or at maxiter
    auto finish_level = [&]() -> bool {
        level_done();
        if (++lev < n_levels) return false;
        lev = 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            double o = __shfl_xor(err, off, 64);
            err = (err > o) ? err : o;
        }
        it++;
        if (err < tol) {
            iters = it;
            return true;
        }
        err = 0.0;
        return it >= maxiter;
    };
    for (;;) {
        load(b1, (lev + 1 < n_levels) ? lev + 1 : 0);
        if (b0.deg > 0) {
            const double ch = ct_step(b0, y, frow, e_q, e_geo);
            err = (err > ch) ? err : ch;
        }
        if (finish_level()) break;
        load(b0, (lev + 1 < n_levels) ? lev + 1 : 0);
        if (b1.deg > 0) {
            const double ch = ct_step(b1, y, frow, e_q, e_geo);
            err = (err > ch) ? err : ch;
        }
        if (finish_level()) break;
    }
    if (LDS_Y) {
        __syncthreads();
        for (int i = lane; i < 2 * N; i += 64) grad_g[(int64_t)z * N * 2 + i] = y[i];
    }
    if (lane == 0) iters_out[z] = iters;
}

// Clough-Tocher value at barycentric b in triangle s (vertex order of `simp`).
__device__ double ct_value(const double* __restrict__ pts, const int32_t* __restrict__ simp, const double* g3,
                           const double* b, const double* f, const double* df) {
    const int v0 = simp[0], v1 = simp[1], v2 = simp[2];
    const double e12x = pts[2 * v1] - pts[2 * v0], e12y = pts[2 * v1 + 1] - pts[2 * v0 + 1];
    const double e23x = pts[2 * v2] - pts[2 * v1], e23y = pts[2 * v2 + 1] - pts[2 * v1 + 1];
    const double e31x = pts[2 * v0] - pts[2 * v2], e31y = pts[2 * v0 + 1] - pts[2 * v2 + 1];
    const double f1 = f[0], f2 = f[1], f3 = f[2];
    const double df12 = +(df[0] * e12x + df[1] * e12y);
    const double df21 = -(df[2] * e12x + df[3] * e12y);
    const double df23 = +(df[2] * e23x + df[3] * e23y);
    const double df32 = -(df[4] * e23x + df[5] * e23y);
    const double df31 = +(df[4] * e31x + df[5] * e31y);
    const double df13 = -(df[0] * e31x + df[1] * e31y);
    const double c3000 = f1;
    const double c2100 = (df12 + 3 * c3000) / 3;
    const double c2010 = (df13 + 3 * c3000) / 3;
    const double c0300 = f2;
    const double c1200 = (df21 + 3 * c0300) / 3;
    const double c0210 = (df23 + 3 * c0300) / 3;
    const double c0030 = f3;
    const double c1020 = (df31 + 3 * c0030) / 3;
    const double c0120 = (df32 + 3 * c0030) / 3;
    const double c2001 = (c2100 + c2010 + c3000) / 3;
    const double c0201 = (c1200 + c0300 + c0210) / 3;
    const double c0021 = (c1020 + c0120 + c0030) / 3;
    const double c0111 = (g3[0] * (-c0300 + 3 * c0210 - 3 * c0120 + c0030) + (-c0300 + 2 * c0210 - c0120 + c0021 + c0201)) / 2;
    const double c1011 = (g3[1] * (-c0030 + 3 * c1020 - 3 * c2010 + c3000) + (-c0030 + 2 * c1020 - c2010 + c2001 + c0021)) / 2;
    const double c1101 = (g3[2] * (-c3000 + 3 * c2100 - 3 * c1200 + c0300) + (-c3000 + 2 * c2100 - c1200 + c2001 + c0201)) / 2;
    const double c1002 = (c1101 + c1011 + c2001) / 3;
    const double c0102 = (c1101 + c0111 + c0201) / 3;
    const double c0012 = (c1011 + c0111 + c0021) / 3;
    const double c0003 = (c1002 + c0102 + c0012) / 3;
    double minval = b[0];
    for (int k = 0; k < 3; k++)
        if (b[k] < minval) minval = b[k];
    const double b1 = b[0] - minval, b2 = b[1] - minval, b3 = b[2] - minval, b4 = 3 * minval;
    return (b1 * b1 * b1 * c3000 + 3 * b1 * b1 * b2 * c2100 + 3 * b1 * b1 * b3 * c2010 + 3 * b1 * b1 * b4 * c2001 +
            3 * b1 * b2 * b2 * c1200 + 6 * b1 * b2 * b4 * c1101 + 3 * b1 * b3 * b3 * c1020 + 6 * b1 * b3 * b4 * c1011 +
            3 * b1 * b4 * b4 * c1002 + b2 * b2 * b2 * c0300 + 3 * b2 * b2 * b3 * c0210 + 3 * b2 * b2 * b4 * c0201 +
            3 * b2 * b3 * b3 * c0120 + 6 * b2 * b3 * b4 * c0111 + 3 * b2 * b4 * b4 * c0102 + b3 * b3 * b3 * c0030 +
            3 * b3 * b3 * b4 * c0021 + 3 * b3 * b4 * b4 * c0012 + b4 * b4 * b4 * c0003);
}

// One thread per (regular-grid point q, slice z); writes D[i][z][j] for q = i*G + j.
__global__ void __launch_bounds__(256) k_ct_eval(const double* __restrict__ dens, int N, int Z, int G,
                                                 const double* __restrict__ pts, const int32_t* __restrict__ simplices,
                                                 const double* __restrict__ sg, const int32_t* __restrict__ qs,
                                                 const double* __restrict__ qb, const int32_t* __restrict__ qn,
                                                 const double* __restrict__ grad, double* __restrict__ Dm,
                                                 int32_t* __restrict__ bad) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (q >= (int64_t)G * G) return;
    const int s = qs[q];
    double w = __builtin_nan("");
    if (s >= 0) {
        const int32_t* sp = simplices + 3 * s;
        double f[3], df[6], b[3], g3[3];
        for (int k = 0; k < 3; k++) {
            const int v = sp[k];
            f[k] = dens[(int64_t)v * Z + z];
            df[2 * k] = grad[((int64_t)z * N + v) * 2];
            df[2 * k + 1] = grad[((int64_t)z * N + v) * 2 + 1];
            b[k] = qb[3 * q + k];
            g3[k] = sg[3 * s + k];
        }
        w = ct_value(pts, sp, g3, b, f, df);
    }
    if (isnan(w) || w < 0.0) {
        w = dens[(int64_t)qn[q] * Z + z];
        if (isnan(w) || w < 0.0) atomicAdd(bad, 1);
    }
    const int64_t i = q / G, j = q % G;
    Dm[(i * Z + z) * G + j] = w;
}

// C (m x n) = A (m x k) * B (k x n), row-major FP64; 64x64 tiles, 256 threads x (4x4) outputs.
constexpr int GT = 64, GK = 16;
__global__ void __launch_bounds__(256) k_gemm_f64(const double* __restrict__ A, const double* __restrict__ B,
                                                  double* __restrict__ C, int m, int n, int k) {
    __shared__ double sA[GK][GT + 1];
    __shared__ double sB[GK][GT + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.y * GT, col0 = blockIdx.x * GT;
    double acc[4][4] = {{0}};
    for (int k0 = 0; k0 < k; k0 += GK) {
        for (int e = threadIdx.x; e < GT * GK; e += 256) {
            int r = e / GK, kk = e % GK;
            int gr = row0 + r, gk = k0 + kk;
            sA[kk][r] = (gr < m && gk < k) ? A[(int64_t)gr * k + gk] : 0.0;
            int kk2 = e / GT, c = e % GT;
            int gk2 = k0 + kk2, gc = col0 + c;
            sB[kk2][c] = (gk2 < k && gc < n) ? B[(int64_t)gk2 * n + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk++) {
            double a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = sA[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) bb[j] = sB[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int r = row0 + ty * 4 + i;
        if (r >= m) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int c = col0 + tx * 4 + j;
            if (c < n) C[(int64_t)r * n + c] = acc[i][j];
        }
    }
}

// dens[p][z] = sum_c w[p][c] * Y[i_c][z][j_c]
__global__ void __launch_bounds__(256) k_back_interp(const double* __restrict__ Y, int N, int Z, int G,
                                                     const int32_t* __restrict__ bidx, const double* __restrict__ bw,
                                                     double* __restrict__ dens) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * Z) return;
    const int64_t p = t / Z;
    const int z = (int)(t % Z);
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int64_t q = bidx[4 * p + c];
        const int64_t i = q / G, j = q % G;
        v += bw[4 * p + c] * Y[(i * Z + z) * G + j];
    }
    dens[p * Z + z] = v;
}

__global__ void k_zero_i32(int32_t* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace hd

using namespace hd;

extern "C" int hd_gemm_f64(const double* A, const double* B, double* C, int m, int n, int k, void* stream) {
    if (m < 1 || n < 1 || k < 1 || !A || !B || !C) return set_error(HD_ERR_ARG, "hd_gemm_f64: bad arguments");
    dim3 grid((unsigned)cdiv(n, GT), (unsigned)cdiv(m, GT));
    hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, (hipStream_t)stream, A, B, C, m, n, k);
    return check_launch("k_gemm_f64");
}

namespace hd {
// wT[k, j] = W[j, k] = dm[k - j + G - 1]: the transposed Gaussian-CDF weight matrix of
// utils._get_rolled_weights (utils.py:464-488) from its 2G-1 distinct values
__global__ void k_rolled_weights_T(const double* __restrict__ dm, int G, double* __restrict__ wT) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G * G) return;
    const int k = (int)(t / G), j = (int)(t % G);
    wT[t] = dm[k - j + G - 1];
}
}  // namespace hd

extern "C" int hd_scatter_weights_T(const double* dm, int G, double* wT, void* stream) {
    if (!dm || !wT || G < 1) return set_error(HD_ERR_ARG, "hd_scatter_weights_T: bad arguments");
    hipLaunchKernelGGL(k_rolled_weights_T, dim3((unsigned)cdiv((int64_t)G * G, 256)), dim3(256), 0,
                       (hipStream_t)stream, dm, G, wT);
    return check_launch("k_rolled_weights_T");
}

extern "C" size_t hd_scatter_workspace_size(int M, int Q, int Z, int G) {
    size_t N = (size_t)M * Q;
    const size_t ecap = 6 * N;
    return align256(N * Z * 2 * 8) + 2 * align256((size_t)G * G * Z * 8) + align256(N * Z * 8) +   // grad, D, Y, densT
           align256((N + 1) * 4) + align256(N * 4 * 8) + align256(ecap * 4) + align256(ecap * 3 * 8) +
           align256(N * 4) + align256(CT_D * N * 4) + 3 * align256(CT_D * N * 8) +   // padded geometry
           align256((size_t)Z * CT_D * N * 8) + 256;                              // 6 (f1 - f2) per slice
}

extern "C" int hd_scatter(double* dens, int M, int Q, int Z, int G, const double* pts, const int32_t* nbr_ptr,
                          const int32_t* nbr_idx, const int32_t* levels_ptr, const int32_t* levels_idx, int n_levels,
                          const int32_t* simplices, const double* simplex_g, const int32_t* q_simplex,
                          const double* q_bary, const int32_t* q_nearest, const double* P, const int32_t* back_idx,
                          const double* back_w, double tol, int maxiter, void* workspace, size_t ws_bytes,
                          int32_t* iters_out, void* stream) {
    if (M < 2 || Q < 2 || Z < 1 || G < 2 || n_levels < 1 || !dens || !pts || !nbr_ptr || !nbr_idx || !levels_ptr ||
        !levels_idx || !simplices || !simplex_g || !q_simplex || !q_bary || !q_nearest || !P || !back_idx || !back_w ||
        !iters_out)
        return set_error(HD_ERR_ARG, "hd_scatter: bad arguments");
    if ((int64_t)G * G > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_scatter: G too large");
    size_t need = hd_scatter_workspace_size(M, Q, Z, G);
    if (!workspace || ws_bytes < need) return set_error(HD_ERR_SIZE, "hd_scatter: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    const int N = M * Q;
    char* p = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    double* grad = (double*)p;  p += align256((size_t)N * Z * 2 * 8);
    double* Dm = (double*)p;    p += align256((size_t)G * G * Z * 8);
    double* Y = (double*)p;     p += align256((size_t)G * G * Z * 8);
    double* densT = (double*)p; p += align256((size_t)N * Z * 8);
    const int64_t ecap = 6 * (int64_t)N;
    int32_t* slot_eoff = (int32_t*)p;  p += align256((size_t)(N + 1) * 4);
    double* slot_q = (double*)p;       p += align256((size_t)N * 4 * 8);
    int32_t* e_q = (int32_t*)p;        p += align256((size_t)ecap * 4);
    double* e_geo = (double*)p;        p += align256((size_t)ecap * 3 * 8);
    int32_t* s_deg = (int32_t*)p;      p += align256((size_t)N * 4);
    int32_t* pq = (int32_t*)p;         p += align256((size_t)CT_D * N * 4);
    double* pex = (double*)p;          p += align256((size_t)CT_D * N * 8);
    double* pey = (double*)p;          p += align256((size_t)CT_D * N * 8);
    double* pl3 = (double*)p;          p += align256((size_t)CT_D * N * 8);
    double* fdiff = (double*)p;
    int32_t* bad = iters_out + Z;   // iters_out has Z + 1 entries

    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, bad, 1);
    hipLaunchKernelGGL(k_ct_slots, dim3(1), dim3(1024), 0, st, N, nbr_ptr, levels_idx, slot_eoff);
    hipLaunchKernelGGL(k_ct_geom, dim3((unsigned)cdiv(N, 256)), dim3(256), 0, st, N, pts, nbr_ptr, nbr_idx, levels_idx,
                       slot_eoff, ecap, slot_q, e_q, e_geo, bad);
    hipLaunchKernelGGL(k_transpose_nz, dim3((unsigned)cdiv((int64_t)N * Z, 256)), dim3(256), 0, st, dens, N, Z, densT);
    hipLaunchKernelGGL(k_ct_pad, dim3((unsigned)cdiv(N, 256)), dim3(256), 0, st, N, slot_eoff, e_q, e_geo, s_deg, pq,
                       pex, pey, pl3);
    hipLaunchKernelGGL(k_ct_fdiff, dim3((unsigned)cdiv((int64_t)N * Z, 256)), dim3(256), 0, st, densT, N, Z,
                       levels_idx, s_deg, pq, fdiff);
    int rc0 = check_launch("k_ct_fdiff");
    if (rc0) return rc0;
    // solution vector (and the level pointers) in LDS when they fit
    int dev = 0;
    hipError_t he = hipGetDevice(&dev);
    int max_lds = 0;
    if (he == hipSuccess) he = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (he != hipSuccess) return set_error(HD_ERR_HIP, "hd_scatter: device query: %s", hipGetErrorString(he));
    const size_t lptr_bytes = (size_t)(n_levels + 1) * sizeof(int32_t);
    const size_t y_bytes = (size_t)N * 2 * sizeof(double) + lptr_bytes;
    if (lptr_bytes > (size_t)max_lds) return set_error(HD_ERR_SIZE, "hd_scatter: %d levels exceed the LDS", n_levels);
    if (y_bytes <= (size_t)max_lds) {
        if (y_bytes > 65536) {
            he = hipFuncSetAttribute((const void*)k_ct_gradients<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)y_bytes);
            if (he != hipSuccess) return set_error(HD_ERR_HIP, "hd_scatter: LDS attribute: %s", hipGetErrorString(he));
        }
        hipLaunchKernelGGL(k_ct_gradients<true>, dim3(Z), dim3(64), y_bytes, st, densT, N, Z, levels_idx, slot_eoff,
                           slot_q, e_q, e_geo, s_deg, pq, pex, pey, pl3, fdiff, levels_ptr, n_levels, tol, maxiter,
                           grad, iters_out);
    } else {
        if (lptr_bytes > 65536) {
            he = hipFuncSetAttribute((const void*)k_ct_gradients<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lptr_bytes);
            if (he != hipSuccess) return set_error(HD_ERR_HIP, "hd_scatter: LDS attribute: %s", hipGetErrorString(he));
        }
        hipLaunchKernelGGL(k_ct_gradients<false>, dim3(Z), dim3(64), lptr_bytes, st, densT, N, Z, levels_idx, slot_eoff,
                           slot_q, e_q, e_geo, s_deg, pq, pex, pey, pl3, fdiff, levels_ptr, n_levels, tol, maxiter,
                           grad, iters_out);
    }
    int rc = check_launch("k_ct_gradients");
    if (rc) return rc;
    dim3 g1((unsigned)cdiv((int64_t)G * G, 256), Z);
    hipLaunchKernelGGL(k_ct_eval, g1, dim3(256), 0, st, dens, N, Z, G, pts, simplices, simplex_g, q_simplex, q_bary,
                       q_nearest, grad, Dm, bad);
    rc = check_launch("k_ct_eval");
    if (rc) return rc;
    // Y (G x Z*G) = P (G x G) * D (G x Z*G)
    rc = hd_gemm_f64(P, Dm, Y, G, Z * G, G, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_back_interp, dim3((unsigned)cdiv((int64_t)N * Z, 256)), dim3(256), 0, st, Y, N, Z, G, back_idx,
                       back_w, dens);
    return check_launch("k_back_interp");
}
