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
// Adjacency in slot order, padded to CT_SLOT entries per slot: entry k < CT_SLOT
// of slot t is at t CT_SLOT + k (neighbour -1 past the degree); the entries of a
// vertex with more neighbours continue in an overflow area after the N CT_SLOT
// padded ones, at slot_ovf[t] .. slot_ovf[t+1] (exclusive prefix of the
// surplus degrees, made here by one workgroup).
constexpr int CT_SLOT = 8;

__global__ void __launch_bounds__(1024) k_ct_slots(int N, const int32_t* __restrict__ nptr,
                                                   const int32_t* __restrict__ lidx, int32_t* __restrict__ slot_ovf) {
    __shared__ int32_t s_part[1024];
    const int tid = threadIdx.x;
    int32_t carry = 0;
    for (int base = 0; base < N; base += 1024) {
        const int t = base + tid;
        int32_t d = (t < N) ? nptr[lidx[t] + 1] - nptr[lidx[t]] - CT_SLOT : 0;
        d = d > 0 ? d : 0;
        s_part[tid] = d;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int32_t v = (tid >= off) ? s_part[tid - off] : 0;
            __syncthreads();
            s_part[tid] += v;
            __syncthreads();
        }
        if (t < N) slot_ovf[t] = carry + s_part[tid] - d;
        carry += s_part[1023];
        __syncthreads();
    }
    if (tid == 0) slot_ovf[N] = carry;
}

// Index of adjacency entry k of slot t in the padded layout.
__device__ __forceinline__ int64_t ct_entry(int64_t N, int t, int k, const int32_t* __restrict__ slot_ovf) {
    return (k < CT_SLOT) ? (int64_t)t * CT_SLOT + k : N * CT_SLOT + slot_ovf[t] + (k - CT_SLOT);
}

// Geometry of every slot: the vertex record {p, overflow begin, overflow end} and
// {Q0, Q1, Q3, det, 1/det}, the 2x2 system summed over the neighbours in order,
// and per adjacency entry the neighbour and {ex, ey, L^3, 1/L^3} -- the same
// expressions in the same order as the sequential sweep evaluates them.  Pads:
// neighbour -1.  Overflow entries beyond n_ovf are reported through *bad.
__global__ void __launch_bounds__(256) k_ct_geom(int N, const double* __restrict__ pts, const int32_t* __restrict__ nptr,
                                                 const int32_t* __restrict__ nidx, const int32_t* __restrict__ lidx,
                                                 const int32_t* __restrict__ slot_ovf, int n_ovf,
                                                 int4* __restrict__ slot_i, double* __restrict__ slot_v,
                                                 int32_t* __restrict__ pe_q, double4* __restrict__ pe_geo,
                                                 int32_t* __restrict__ bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    if (slot_ovf[N] > n_ovf) {
        if (t == 0) atomicOr(bad, 1 << 20);
        return;
    }
    const int p = lidx[t];
    const double px = pts[2 * p], py = pts[2 * p + 1];
    double Q0 = 0, Q1 = 0, Q3 = 0;
    const int deg = nptr[p + 1] - nptr[p];
    for (int k = 0; k < deg; k++) {
        const int q = nidx[nptr[p] + k];
        const double ex = pts[2 * q] - px;
        const double ey = pts[2 * q + 1] - py;
        const double L = sqrt(ex * ex + ey * ey);
        const double L3 = L * L * L;
        Q0 += 4 * ex * ex / L3;
        Q1 += 4 * ex * ey / L3;
        Q3 += 4 * ey * ey / L3;
        const int64_t e = ct_entry(N, t, k, slot_ovf);
        pe_q[e] = q;
        pe_geo[e] = make_double4(ex, ey, L3, 1.0 / L3);
    }
    for (int k = deg; k < CT_SLOT; k++) {
        pe_q[(int64_t)t * CT_SLOT + k] = -1;
        pe_geo[(int64_t)t * CT_SLOT + k] = make_double4(0.0, 0.0, 1.0, 1.0);
    }
    const double Q2 = Q1;
    const double det = Q0 * Q3 - Q1 * Q2;
    slot_i[t] = make_int4(p, slot_ovf[t], slot_ovf[t + 1], 0);
    double* sv = slot_v + 6 * (int64_t)t;
    sv[0] = Q0;
    sv[1] = Q1;
    sv[2] = Q3;
    sv[3] = det;
    sv[4] = 1.0 / det;
    sv[5] = 0.0;
}

// dens (N, Z) -> densT (Z, N): each slice contiguous
__global__ void __launch_bounds__(256) k_transpose_nz(const double* __restrict__ a, int N, int Z, double* __restrict__ b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * Z) return;
    const int64_t p = t / Z, z = t % Z;
    b[z * N + p] = a[t];
}

// The per-slice constant of every adjacency entry, 6 (f1 - f2) (the sweep's first
// term, the same operation on the same values), A[z][entry]; 0 for pads.
__global__ void __launch_bounds__(256) k_ct_fdiff(const double* __restrict__ densT, int N, int Z, int64_t ecap,
                                                  const int4* __restrict__ slot_i, const int32_t* __restrict__ pe_q,
                                                  double* __restrict__ A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)N * Z) return;
    const int z = (int)(i / N), t = (int)(i % N);
    const double* frow = densT + (int64_t)z * N;
    const int4 si = slot_i[t];
    const double f1 = frow[si.x];
    double* Az = A + (int64_t)z * ecap;
    for (int k = 0; k < CT_SLOT; k++) {
        const int64_t e = (int64_t)t * CT_SLOT + k;
        const int q = pe_q[e];
        Az[e] = (q >= 0) ? 6 * (f1 - frow[q]) : 0.0;
    }
    for (int64_t e = (int64_t)N * CT_SLOT + si.y; e < (int64_t)N * CT_SLOT + si.z; e++)
        Az[e] = 6 * (f1 - frow[pe_q[e]]);
}

// Markstein's final division step (P. Markstein, IBM J. Res. Dev. 34 (1990) 111):
// for b > 0 and rb = RN(1/b), q0 = RN(a rb), the remainder a - b q0 exactly by one
// fma, q = RN(q0 + remainder rb) is a / b correctly rounded while |q0| stays in
// [2^-900, 2^900] (so neither the remainder nor the correction leaves the normal
// range), and for a zero a it gives q0, the zero a / b gives.  ct_div_ok tells
// whether one of the two holds.
__device__ __forceinline__ double ct_div_fast(double a, double b, double rb, double q0) {
    const double r = fma(b, q0, -a);   // -(a - b q0): +0 when exact, so a zero a keeps its sign in q0
    return fma(-r, rb, q0);
}
__device__ __forceinline__ bool ct_div_ok(double a, double q0) {
    const double aq = fabs(q0);
    return ((aq >= 0x1p-900) & (aq <= 0x1p+900)) | (a == 0.0);   // no short circuit: no branches
}

// scipy's two terms of one edge, ((6 (f1 - f2) - 2 df2) ex) / L^3 and
// ((6 (f1 - f2) - 2 df2) ey) / L^3 with df2 = -ex y[q].x - ey y[q].y; exactly
// rounded divisions (the IEEE division itself where the fast step's range test
// fails).
__device__ __forceinline__ double2 ct_terms(double a, const double4 g, const double2 yq) {
    const double df2 = -g.x * yq.x - g.y * yq.y;
    const double u = a - 2 * df2;
    const double n0 = u * g.x, n1 = u * g.y;
    const double q0 = n0 * g.w, q1 = n1 * g.w;
    double2 r;
    r.x = ct_div_fast(n0, g.z, g.w, q0);
    r.y = ct_div_fast(n1, g.z, g.w, q1);
    if (!(ct_div_ok(n0, q0) & ct_div_ok(n1, q1))) {   // rare: exactly rounded either way
        r.x = n0 / g.z;
        r.y = n1 / g.z;
    }
    return r;
}

// LDS writes complete, then the workgroup barrier; the loads in flight are not
// waited for (the inline asm is also a compiler memory barrier, so no LDS access
// moves across it).
__device__ __forceinline__ void ct_barrier_lds() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Gauss-Seidel sweeps of one slice by one workgroup.  Within a level, a group of
// CT_SLOT lanes per vertex: lane k computes the terms of adjacency entry k (the
// inputs loaded one level ahead) into LDS, -0.0 for a pad (x + -0.0 = x exactly,
// so a pad changes no sum); the group's first lane reads them back (the lanes
// are in its own wavefront, whose LDS operations complete in order, so no
// barrier), adds them in the neighbour order as the sequential loop does, adds
// the terms of any overflow entries itself, solves the 2x2 system and writes the
// vertex's gradient.  One barrier per level.  Arithmetic and order are scipy's
// term by term, so the gradients equal a sequential sweep's bit for bit.  The
// convergence measure, max over vertices of max|y_old + r| / max(1, max|r|) with
// y_new = -r, is taken at the end of each sweep from the previous sweep's
// gradients kept in grad_g (the same operations on the same values).  The
// workgroup holds CT_SLOT lanes for every vertex of the widest level; the
// kernel checks that and refuses otherwise.
constexpr int CT_MAX_THREADS = 1024;
constexpr int CT_SWEEP_UNROLL = 4;   // vertices per thread in flight in the end-of-sweep pass

template <bool LDS_Y>
__global__ void __launch_bounds__(CT_MAX_THREADS) k_ct_gradients(
    int N, int Z, const int4* __restrict__ slot_i, const double* __restrict__ slot_v, const int32_t* __restrict__ pe_q,
    const double4* __restrict__ pe_geo, const double* __restrict__ A, int64_t ecap, const int32_t* __restrict__ lptr,
    int n_levels, double tol, int maxiter, double* __restrict__ ybuf, double* __restrict__ grad_g,
    int32_t* __restrict__ iters_out) {
    extern __shared__ double4 s_dyn[];
    __shared__ double s_err[CT_MAX_THREADS / 64];
    __shared__ int s_bad;
    const int z = blockIdx.x;
    const int nthr = blockDim.x, nw = nthr >> 6;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int j = tid / CT_SLOT, k = tid % CT_SLOT;
    const double* Az = A + (int64_t)z * ecap;
    double2* gz = (double2*)(grad_g + (int64_t)z * N * 2);
    // LDS: [solution N double2 (if LDS_Y)] [terms nthr double2] [level slot pointers n+1]
    double2* y = LDS_Y ? (double2*)s_dyn : (double2*)(ybuf + (int64_t)z * N * 2);
    double2* T = LDS_Y ? (double2*)s_dyn + N : (double2*)s_dyn;
    int32_t* s_lptr = (int32_t*)(T + nthr);
    for (int i = tid; i < N; i += nthr) {
        y[i] = make_double2(0.0, 0.0);
        gz[i] = make_double2(0.0, 0.0);   // the previous sweep's gradients (read back by this thread only)
    }
    if (tid == 0) s_bad = 0;
    for (int i = tid; i <= n_levels; i += nthr) s_lptr[i] = lptr[i];
    __syncthreads();
    for (int i = tid; i < n_levels; i += nthr)   // a level wider than the workgroup: refuse
        if ((s_lptr[i + 1] - s_lptr[i]) * CT_SLOT > nthr) s_bad = 1;
    __syncthreads();
    if (s_bad) {
        if (tid == 0) {
            atomicOr(iters_out + Z, 1 << 21);
            iters_out[z] = 0;
        }
        return;
    }
    const uint32_t elast = (uint32_t)N * CT_SLOT - 1;
    // this lane's adjacency entry and vertex, one level ahead; every load issued on
    // every lane (indices clamped into the arrays), so no load sits under a branch
    // and a use waits for its own loads only.  Byte offsets are 32-bit (the host
    // checks the arrays' sizes), so the loads take a scalar base and a 32-bit offset.
    int cq;
    double4 cg;
    double ca;
    int4 vi;
    double vQ0, vQ1, vQ3, vdet, vrdet;
    auto load_level = [&](int l0, bool vert) {
        if (!vert) {
            const uint32_t e = min((uint32_t)((l0 + j) * CT_SLOT + k), elast);
            cq = *(const int32_t*)((const char*)pe_q + (e << 2));
            cg = *(const double4*)((const char*)pe_geo + (e << 5));
            ca = *(const double*)((const char*)Az + (e << 3));
        } else {
            const uint32_t t = min((uint32_t)(l0 + j), (uint32_t)(N - 1));
            vi = *(const int4*)((const char*)slot_i + (t << 4));
            const double* sv = (const double*)((const char*)slot_v + t * 48u);
            vQ0 = sv[0];
            vQ1 = sv[1];
            vQ3 = sv[2];
            vdet = sv[3];
            vrdet = sv[4];
        }
    };
    int lb0 = s_lptr[0], lb1 = s_lptr[1];
    int nb0 = s_lptr[n_levels > 1 ? 1 : 0], nb1 = s_lptr[n_levels > 1 ? 2 : 1];
    load_level(lb0, false);
    load_level(lb0, true);
    int lev = 0, it = 0, iters = 0;
    for (;;) {
        const bool act = j < lb1 - lb0;
        // the entry's terms
        if (act) {
            const double2 yq = y[cq >= 0 ? cq : 0];
            double2 tt = ct_terms(ca, cg, yq);
            if (cq < 0) tt = make_double2(-0.0, -0.0);
            T[tid] = tt;
        }
        load_level(nb0, false);   // the next level's entries
        asm volatile("" ::: "memory");
        // the group's first lane: in-order sums, overflow entries, solve
        if (act && k == 0) {
            double2 tk[CT_SLOT];
#pragma unroll
            for (int i = 0; i < CT_SLOT; i++) tk[i] = T[tid + i];
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int i = 0; i < CT_SLOT; i++) {
                s0 += tk[i].x;
                s1 += tk[i].y;
            }
            for (int64_t e = (int64_t)N * CT_SLOT + vi.y; e < (int64_t)N * CT_SLOT + vi.z; e++) {   // rare
                const int q = pe_q[e];
                const double2 tt = ct_terms(Az[e], pe_geo[e], y[q]);
                s0 += tt.x;
                s1 += tt.y;
            }
            const double Q2 = vQ1;
            const double n0 = vQ3 * s0 - vQ1 * s1;
            const double n1 = -Q2 * s0 + vQ0 * s1;
            const double q0 = n0 * vrdet, q1 = n1 * vrdet;
            double r0 = ct_div_fast(n0, vdet, vrdet, q0);
            double r1 = ct_div_fast(n1, vdet, vrdet, q1);
            if (!((vdet > 0.0) & ct_div_ok(n0, q0) & ct_div_ok(n1, q1))) {   // rare: exactly rounded either way
                r0 = n0 / vdet;
                r1 = n1 / vdet;
            }
            y[vi.x] = make_double2(-r0, -r1);
        }
        load_level(nb0, true);   // the next level's vertices
        {
            const int ln2 = (lev + 2 < n_levels) ? lev + 2 : lev + 2 - n_levels;   // wraps into the next sweep
            lb0 = nb0;
            lb1 = nb1;
            nb0 = s_lptr[ln2 < n_levels ? ln2 : 0];
            nb1 = s_lptr[(ln2 < n_levels ? ln2 : 0) + 1];
        }
        if (LDS_Y) ct_barrier_lds();
        else __syncthreads();
        if (++lev < n_levels) continue;
        // end of a sweep: the largest relative change over every vertex, against the
        // previous sweep's gradients, which the new ones then replace
        lev = 0;
        it++;
        double err = 0.0;
        for (int base = 0; base < N; base += CT_SWEEP_UNROLL * nthr) {
            double2 yn[CT_SWEEP_UNROLL], yo[CT_SWEEP_UNROLL];
#pragma unroll
            for (int u = 0; u < CT_SWEEP_UNROLL; u++) {   // loads in flight together (index clamped)
                const int i = min(base + u * nthr + tid, N - 1);
                yn[u] = y[i];
                yo[u] = gz[i];
            }
#pragma unroll
            for (int u = 0; u < CT_SWEEP_UNROLL; u++) {
                const int i = base + u * nthr + tid;
                if (i < N) {
                    const double a0 = fabs(yo[u].x - yn[u].x), a1 = fabs(yo[u].y - yn[u].y);   // |y_old + r|, r = -y_new
                    const double m0 = fabs(yn[u].x), m1 = fabs(yn[u].y);
                    const double num = (a0 > a1) ? a0 : a1;
                    const double mr = (m0 > m1) ? m0 : m1;
                    const double den = (1.0 > mr) ? 1.0 : mr;
                    const double ch = num / den;
                    err = (err > ch) ? err : ch;
                    gz[i] = yn[u];
                }
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            double o = __shfl_xor(err, off, 64);
            err = (err > o) ? err : o;
        }
        if (lane == 0) s_err[w] = err;
        __syncthreads();
        double m = s_err[0];
        for (int i = 1; i < nw; i++) m = (m > s_err[i]) ? m : s_err[i];
        __syncthreads();   // s_err is read by every wave before the next sweep can write it
        if (m < tol) {
            iters = it;
            break;
        }
        if (it >= maxiter) break;
    }
    if (tid == 0) iters_out[z] = iters;
}

// The sweeps pipelined: sweep s + 1 starts on its early levels while sweep s is still on
// its late ones.  (Vertex i, sweep s) runs at step l(i) + s P, P one more than the largest
// level difference across an edge: a vertex then reads its lower neighbours after their
// update of the same sweep and its higher neighbours after their update of the previous
// sweep and before their update of this one -- exactly the values the sequential sweep
// reads, so the result is bit for bit the same.  A step runs the levels l = t - s P of the
// sweeps in flight (at most ceil(levels / P)), one group of four lanes per vertex (two
// entries each; otherwise as in k_ct_gradients).  Each update also writes its sweep's value to a ring of K sweeps in
// global memory (y_hist) and raises the sweep's largest relative change (an LDS atomic max
// on the non-negative doubles' bits); when a sweep's last level is done its change is
// compared with tol, and on convergence (or at maxiter) the gradients are that sweep's ring
// slot: the later sweeps in flight have overwritten some of the solution in LDS, but they
// write other ring slots, and slot s is reused only by sweep s + K, which starts after the
// check (K P > levels + P; the change slots likewise).
// Step tables per residue r = t mod P (levels r, r + P, ...: their first slots, prefix
// widths, and the map from a group to its level and place in it) are built in LDS at the start.  The workgroup holds four lanes for every vertex of
// the widest step; the host checks that and the LDS before choosing this kernel.
constexpr int CT_MAX_K = 16;
constexpr int CT_PIPE_LANES = 4;   // lanes per vertex in the pipelined sweeps (two entries each)

__global__ void __launch_bounds__(CT_MAX_THREADS) k_ct_gradients_pipe(
    int N, int Z, const int4* __restrict__ slot_i, const double* __restrict__ slot_v, const int32_t* __restrict__ pe_q,
    const double4* __restrict__ pe_geo, const double* __restrict__ A, int64_t ecap, const int32_t* __restrict__ lptr,
    int n_levels, int P, int K, int W, double tol, int maxiter, double* __restrict__ y_hist,
    double* __restrict__ grad_g, int32_t* __restrict__ iters_out) {
    extern __shared__ double4 s_dyn[];
    __shared__ unsigned long long s_err[CT_MAX_K];
    __shared__ int s_bad;
    const int z = blockIdx.x;
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;
    const int j = tid / CT_PIPE_LANES, k = tid % CT_PIPE_LANES;   // lane k holds entries k and k + 4
    const double* Az = A + (int64_t)z * ecap;
    double2* hz = (double2*)(y_hist + (int64_t)z * K * N * 2);
    const int KM = (n_levels + P - 1) / P;   // levels per residue, at most
    // LDS: [solution N double2] [terms 8 per group double2] [level slot pointers n+1] [per residue:
    // level count P, first slots P KM, prefix widths P (KM + 1), step map P W (16-bit)]
    double2* y = (double2*)s_dyn;
    double2* T = y + N;   // the terms of group g's entries at T[g CT_SLOT .. g CT_SLOT + 7]
    int32_t* s_lptr = (int32_t*)(T + (nthr / CT_PIPE_LANES) * CT_SLOT);
    int32_t* s_cnt = s_lptr + n_levels + 1;
    int32_t* s_start = s_cnt + P;
    int32_t* s_pre = s_start + P * KM;
    uint16_t* s_map = (uint16_t*)(s_pre + P * (KM + 1));   // (level k of the residue) << 12 | index in the level
    for (int i = tid; i < N; i += nthr) y[i] = make_double2(0.0, 0.0);
    for (int i = tid; i <= n_levels; i += nthr) s_lptr[i] = lptr[i];
    if (tid < CT_MAX_K) s_err[tid] = 0ull;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int r = tid; r < P; r += nthr) {
        int c = 0, acc = 0;
        for (int l = r; l < n_levels; l += P, c++) {
            s_start[r * KM + c] = s_lptr[l];
            s_pre[r * (KM + 1) + c] = acc;
            acc += s_lptr[l + 1] - s_lptr[l];
        }
        s_pre[r * (KM + 1) + c] = acc;
        s_cnt[r] = c;
        if (acc * CT_PIPE_LANES > nthr || acc > W) s_bad = 1;   // a step wider than the workgroup: refuse
        else
            for (int l = r, kk = 0; l < n_levels; l += P, kk++)
                for (int i = 0; i < s_lptr[l + 1] - s_lptr[l]; i++)
                    s_map[r * W + s_pre[r * (KM + 1) + kk] + i] = (uint16_t)((kk << 12) | i);
    }
    __syncthreads();
    if (s_bad) {
        if (tid == 0) {
            atomicOr(iters_out + Z, 1 << 21);
            iters_out[z] = 0;
        }
        return;
    }
    // the vertex of this lane's group at step t = q P + r (false: none) and its sweep; the
    // prefix widths of the residue are read together and counted, no dependent search
    auto locate = [&](int q, int r, int& slot, int& sweep) -> bool {
        if (q >= KM - 1 && q < maxiter) {   // every level of every residue in flight: the map directly
            if (j >= s_pre[r * (KM + 1) + s_cnt[r]]) return false;
            const int e = s_map[r * W + j];
            slot = s_start[r * KM + (e >> 12)] + (e & 4095);
            sweep = q - (e >> 12);
            return true;
        }
        const int cnt = s_cnt[r];
        const int khi = min(cnt - 1, q);                 // levels r + k P <= t
        const int klo = max(0, q - maxiter + 1);         // sweeps (t - l) / P < maxiter
        if (klo > khi) return false;
        const int32_t* pre = s_pre + r * (KM + 1);
        const int g = j + pre[klo];
        if (g >= pre[khi + 1]) return false;
        const int e = s_map[r * W + g];
        slot = s_start[r * KM + (e >> 12)] + (e & 4095);
        sweep = q - (e >> 12);
        return true;
    };
    const uint32_t elast = (uint32_t)N * CT_SLOT - 1;
    int cq, cq2;
    double4 cg, cg2;
    double ca, ca2;
    int4 vi;
    double vQ0, vQ1, vQ3, vdet, vrdet;
    auto load_entry = [&](int slot) {
        const uint32_t e = min((uint32_t)(slot * CT_SLOT + k), elast);
        const uint32_t e2 = min(e + CT_PIPE_LANES, elast);
        cq = *(const int32_t*)((const char*)pe_q + (e << 2));
        cq2 = *(const int32_t*)((const char*)pe_q + (e2 << 2));
        cg = *(const double4*)((const char*)pe_geo + (e << 5));
        cg2 = *(const double4*)((const char*)pe_geo + (e2 << 5));
        ca = *(const double*)((const char*)Az + (e << 3));
        ca2 = *(const double*)((const char*)Az + (e2 << 3));
    };
    auto load_vert = [&](int slot) {
        const uint32_t t = min((uint32_t)slot, (uint32_t)(N - 1));
        vi = *(const int4*)((const char*)slot_i + (t << 4));
        const double* sv = (const double*)((const char*)slot_v + t * 48u);
        vQ0 = sv[0];
        vQ1 = sv[1];
        vQ3 = sv[2];
        vdet = sv[3];
        vrdet = sv[4];
    };
    // steps t (current) and t + 1 located; t + 2 is located during step t, after its updates
    int slot = 0, sweep = 0, slot1 = 0, sweep1 = 0;
    int q2 = 2 / P, r2 = 2 % P;   // t + 2 = q2 P + r2
    bool act = locate(0, 0, slot, sweep);
    bool act1 = locate(1 / P, 1 % P, slot1, sweep1);
    load_entry(act ? slot : 0);
    load_vert(act ? slot : 0);
    int iters = 0, fin = 0;
    for (int t = 0;; t++) {
        // the entry's terms
        if (act) {
            const double2 yq = y[cq >= 0 ? cq : 0], yq2 = y[cq2 >= 0 ? cq2 : 0];
            double2 tt = ct_terms(ca, cg, yq), tt2 = ct_terms(ca2, cg2, yq2);
            if (cq < 0) tt = make_double2(-0.0, -0.0);
            if (cq2 < 0) tt2 = make_double2(-0.0, -0.0);
            T[j * CT_SLOT + k] = tt;
            T[j * CT_SLOT + k + CT_PIPE_LANES] = tt2;
        }
        load_entry(act1 ? slot1 : 0);   // the next step's entry
        asm volatile("" ::: "memory");
        // the group's first lane: in-order sums, overflow entries, solve, change, ring
        if (act && k == 0) {
            double2 tk[CT_SLOT];
#pragma unroll
            for (int i = 0; i < CT_SLOT; i++) tk[i] = T[j * CT_SLOT + i];
            const double2 yo = y[vi.x];
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int i = 0; i < CT_SLOT; i++) {
                s0 += tk[i].x;
                s1 += tk[i].y;
            }
            for (int64_t e = (int64_t)N * CT_SLOT + vi.y; e < (int64_t)N * CT_SLOT + vi.z; e++) {   // rare
                const int q = pe_q[e];
                const double2 tt = ct_terms(Az[e], pe_geo[e], y[q]);
                s0 += tt.x;
                s1 += tt.y;
            }
            const double Q2 = vQ1;
            const double n0 = vQ3 * s0 - vQ1 * s1;
            const double n1 = -Q2 * s0 + vQ0 * s1;
            const double q0 = n0 * vrdet, q1 = n1 * vrdet;
            double r0 = ct_div_fast(n0, vdet, vrdet, q0);
            double r1 = ct_div_fast(n1, vdet, vrdet, q1);
            if (!((vdet > 0.0) & ct_div_ok(n0, q0) & ct_div_ok(n1, q1))) {   // rare: exactly rounded either way
                r0 = n0 / vdet;
                r1 = n1 / vdet;
            }
            const double2 yn = make_double2(-r0, -r1);
            y[vi.x] = yn;
            hz[(int64_t)(sweep % K) * N + vi.x] = yn;
            // scipy's change: max(|y_old + r|) / max(1, max|r|)
            const double a0 = fabs(yo.x + r0), a1 = fabs(yo.y + r1);
            const double m0 = fabs(r0), m1 = fabs(r1);
            const double num = (a0 > a1) ? a0 : a1;
            const double mr = (m0 > m1) ? m0 : m1;
            const double den = (1.0 > mr) ? 1.0 : mr;
            const double ch = num / den;
            atomicMax(s_err + sweep % K, (unsigned long long)__double_as_longlong(ch));
        }
        load_vert(act1 ? slot1 : 0);   // the next step's vertex
        int slot2 = 0, sweep2 = 0;
        const bool act2 = locate(q2, r2, slot2, sweep2);
        if (++r2 == P) {
            r2 = 0;
            q2++;
        }
        ct_barrier_lds();
        act = act1;
        slot = slot1;
        sweep = sweep1;
        act1 = act2;
        slot1 = slot2;
        sweep1 = sweep2;
        // a sweep whose last level was this step: its convergence test
        const int tl = t - (n_levels - 1);
        if (tl >= 0 && tl % P == 0) {
            const int sc = tl / P;
            const double m = __longlong_as_double((long long)s_err[sc % K]);
            if (m < tol) {
                iters = sc + 1;
                fin = sc;
                break;
            }
            if (sc + 1 >= maxiter) {
                fin = sc;
                break;
            }
            __syncthreads();   // every thread has read the slot before it is cleared
            if (tid == 0) s_err[sc % K] = 0ull;
        }
    }
    __syncthreads();   // the ring writes of every thread are complete
    const double2* src = hz + (int64_t)(fin % K) * N;
    double2* dst = (double2*)(grad_g + (int64_t)z * N * 2);
    for (int i = tid; i < N; i += nthr) dst[i] = src[i];
    if (tid == 0) iters_out[z] = iters;
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

// ---------------------------------------------------------------------------
// Location of the regular (G x G) query grid in the triangulation, on the device
// (what the reference's CloughTocher2DInterpolator and NearestNDInterpolator look
// up on the host: the containing simplex with its barycentric coordinates, and
// the nearest node).  xi[i G + j] = (mgrid[i], mgrid[j]).
// ---------------------------------------------------------------------------
constexpr double CT_BARY_EPS = 100.0 * 2.220446049250313e-16;   // Qhull's inside test: c >= -eps, c <= 1 + eps

// Barycentric coordinates of (x, y) in simplex s from its affine transform
// T (S, 3, 2): c01 = T[:2] (xy - T[2]), c2 = 1 - (c0 + c1).
__device__ __forceinline__ void ct_bary(const double* __restrict__ T, int s, double x, double y, double& c0,
                                        double& c1, double& c2) {
    const double* t = T + 6 * (int64_t)s;
    const double dx = x - t[4], dy = y - t[5];
    c0 = t[0] * dx + t[1] * dy;
    c1 = t[2] * dx + t[3] * dy;
    c2 = 1.0 - (c0 + c1);
}

// One thread per simplex: every grid point of its bounding box (+-2 nodes for the
// grid's rounding) that passes the inside test takes the lowest such simplex
// (a point on a shared edge may pass for both; the interpolant is continuous there).
__global__ void __launch_bounds__(256) k_ct_raster(const double* __restrict__ pts, const int32_t* __restrict__ simp,
                                                   const double* __restrict__ T, int S,
                                                   const double* __restrict__ mgrid, int G,
                                                   int32_t* __restrict__ q_simplex) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int k = 0; k < 3; k++) {
        const int v = simp[3 * s + k];
        xmin = fmin(xmin, pts[2 * v]);
        xmax = fmax(xmax, pts[2 * v]);
        ymin = fmin(ymin, pts[2 * v + 1]);
        ymax = fmax(ymax, pts[2 * v + 1]);
    }
    const double g0 = mgrid[0], inv = (double)(G - 1) / (mgrid[G - 1] - g0);
    const int i0 = max((int)floor((xmin - g0) * inv) - 2, 0), i1 = min((int)ceil((xmax - g0) * inv) + 2, G - 1);
    const int j0 = max((int)floor((ymin - g0) * inv) - 2, 0), j1 = min((int)ceil((ymax - g0) * inv) + 2, G - 1);
    for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++) {
            double c0, c1, c2;
            ct_bary(T, s, mgrid[i], mgrid[j], c0, c1, c2);
            const double lo = -CT_BARY_EPS, hi = 1.0 + CT_BARY_EPS;
            if (c0 >= lo && c1 >= lo && c2 >= lo && c0 <= hi && c1 <= hi && c2 <= hi)
                atomicMin(q_simplex + (int64_t)i * G + j, s);
        }
}

__global__ void k_ct_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per grid point: -1 where no simplex holds it, else its barycentric coordinates in
// the chosen simplex; and the nearest node (squared distance, the lowest index on a
// tie), the nodes staged through LDS in blocks.
__global__ void __launch_bounds__(256) k_ct_bary_nearest(const double* __restrict__ pts, int N,
                                                         const double* __restrict__ T,
                                                         const double* __restrict__ mgrid, int G,
                                                         int32_t* __restrict__ q_simplex, double* __restrict__ q_bary,
                                                         int32_t* __restrict__ q_nearest) {
    __shared__ double s_p[2 * 256];
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = g < (int64_t)G * G;
    const double x = in ? mgrid[g / G] : 0.0, y = in ? mgrid[g % G] : 0.0;
    if (in) {
        const int s = q_simplex[g];
        if (s == 0x7fffffff) {
            q_simplex[g] = -1;
            q_bary[3 * g] = q_bary[3 * g + 1] = q_bary[3 * g + 2] = 0.0;
        } else {
            double c0, c1, c2;
            ct_bary(T, s, x, y, c0, c1, c2);
            q_bary[3 * g] = c0;
            q_bary[3 * g + 1] = c1;
            q_bary[3 * g + 2] = c2;
        }
    }
    double best = 1e308;
    int bi = 0;
    for (int base = 0; base < N; base += 256) {
        __syncthreads();
        const int v = base + threadIdx.x;
        if (v < N) {
            s_p[2 * threadIdx.x] = pts[2 * v];
            s_p[2 * threadIdx.x + 1] = pts[2 * v + 1];
        }
        __syncthreads();
        const int nb = min(256, N - base);
        for (int k = 0; k < nb; k++) {
            const double dx = x - s_p[2 * k], dy = y - s_p[2 * k + 1];
            const double d = dx * dx + dy * dy;
            if (d < best) {
                best = d;
                bi = base + k;
            }
        }
    }
    if (in) q_nearest[g] = bi;
}

// C = A B (row-major, FP64) for the large products: 128 x 128 output tiles, 256 threads with
// 8 x 8 outputs each (64 accumulators: four FMAs per LDS word read, against two for the 64 x 64
// tiles above), K in slices of 8 staged through LDS (A transposed), the next slice's global
// loads issued before the current slice is multiplied.  A thread's 8 columns are two groups of
// 4 at tx*4 and 64 + tx*4, its 8 rows ty*4 and 64 + ty*4, so that a wave's LDS reads and global
// stores are contiguous.
constexpr int GT2 = 128, GK2 = 8;
__global__ void __launch_bounds__(256) k_gemm_f64_big(const double* __restrict__ A, const double* __restrict__ B,
                                                      double* __restrict__ C, int m, int n, int k) {
    __shared__ double sA[GK2][GT2];
    __shared__ double sB[GK2][GT2];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    const int row0 = blockIdx.y * GT2, col0 = blockIdx.x * GT2;
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = 0.0;
    // staging: A tile 128 x 8 (4 words per thread: row ar + 32 q, column ak), B tile 8 x 128
    // (4 words per thread: row bk, column bc + 32 q)
    const int ar = tid / GK2, ak = tid % GK2;
    const int bk = tid / 32, bc = tid % 32;
    double ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gr = row0 + ar + 32 * q, gk = k0 + ak;
            ra[q] = (gr < m && gk < k) ? A[(int64_t)gr * k + gk] : 0.0;
            const int gk2 = k0 + bk, gc = col0 + bc + 32 * q;
            rb[q] = (gk2 < k && gc < n) ? B[(int64_t)gk2 * n + gc] : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < k; k0 += GK2) {
        __syncthreads();   // the previous slice's readers are done
#pragma unroll
        for (int q = 0; q < 4; q++) {
            sA[ak][ar + 32 * q] = ra[q];
            sB[bk][bc + 32 * q] = rb[q];
        }
        __syncthreads();
        if (k0 + GK2 < k) load(k0 + GK2);   // in flight while this slice is multiplied
#pragma unroll
        for (int kk = 0; kk < GK2; kk++) {
            double a[8], b[8];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                a[i] = sA[kk][ty * 4 + i];
                a[4 + i] = sA[kk][64 + ty * 4 + i];
                b[i] = sB[kk][tx * 4 + i];
                b[4 + i] = sB[kk][64 + tx * 4 + i];
            }
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int r = row0 + (i < 4 ? ty * 4 + i : 64 + ty * 4 + i - 4);
        if (r >= m) continue;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = col0 + (j < 4 ? tx * 4 + j : 64 + tx * 4 + j - 4);
            if (c < n) C[(int64_t)r * n + c] = acc[i][j];
        }
    }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace hd

using namespace hd;

extern "C" int hd_gemm_f64(const double* A, const double* B, double* C, int m, int n, int k, void* stream) {
    if (m < 1 || n < 1 || k < 1 || !A || !B || !C) return set_error(HD_ERR_ARG, "hd_gemm_f64: bad arguments");
    if ((int64_t)m * n >= (int64_t)64 * GT2 * GT2) {   // enough 128 x 128 tiles to fill the GPU
        dim3 grid((unsigned)cdiv(n, GT2), (unsigned)cdiv(m, GT2));
        hipLaunchKernelGGL(k_gemm_f64_big, grid, dim3(256), 0, (hipStream_t)stream, A, B, C, m, n, k);
        return check_launch("k_gemm_f64_big");
    }
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

// sweeps kept in the pipelined sweep's ring (0: the sweeps run one after another)
static int ct_ring(int n_levels, int sweep_period) {
    if (sweep_period < 1 || n_levels < 1) return 0;
    const int K = (n_levels + sweep_period) / sweep_period + 2;
    return K <= CT_MAX_K ? K : 0;
}

extern "C" size_t hd_scatter_workspace_size(int M, int Q, int Z, int G, int n_ovf, int n_levels, int sweep_period) {
    size_t N = (size_t)M * Q;
    const size_t ecap = CT_SLOT * N + (size_t)(n_ovf > 0 ? n_ovf : 0);
    const size_t K = (size_t)ct_ring(n_levels, sweep_period);
    return align256(N * Z * 2 * 8) + 2 * align256((size_t)G * G * Z * 8) + align256(N * Z * 8) +   // grad, D, Y, densT
           align256((N + 1) * 4) + align256(N * 16) + align256(N * 6 * 8) +   // slot overflow offsets and records
           align256(ecap * 4) + align256(ecap * 32) +                         // adjacency entries
           align256(N * Z * 2 * 8) +                                          // solution when it exceeds the LDS
           align256((size_t)Z * ecap * 8) +                                   // 6 (f1 - f2) per slice and entry
           align256(K * N * Z * 2 * 8) + 256;                                 // the pipelined sweeps' ring
}

extern "C" int hd_scatter(double* dens, int M, int Q, int Z, int G, const double* pts, const int32_t* nbr_ptr,
                          const int32_t* nbr_idx, const int32_t* levels_ptr, const int32_t* levels_idx, int n_levels,
                          int max_level_width, int n_ovf, int sweep_period, int max_step_width,
                          const int32_t* simplices, const double* simplex_g, const int32_t* q_simplex,
                          const double* q_bary, const int32_t* q_nearest, const double* P, const int32_t* back_idx,
                          const double* back_w, double tol, int maxiter, void* workspace, size_t ws_bytes,
                          int32_t* iters_out, void* stream) {
    if (M < 2 || Q < 2 || Z < 1 || G < 2 || n_levels < 1 || max_level_width < 1 || n_ovf < 0 || maxiter < 1 || !dens || !pts ||
        !nbr_ptr || !nbr_idx || !levels_ptr || !levels_idx || !simplices || !simplex_g || !q_simplex || !q_bary ||
        !q_nearest || !P || !back_idx || !back_w || !iters_out)
        return set_error(HD_ERR_ARG, "hd_scatter: bad arguments");
    if ((int64_t)G * G > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_scatter: G too large");
    size_t need = hd_scatter_workspace_size(M, Q, Z, G, n_ovf, n_levels, sweep_period);
    if (!workspace || ws_bytes < need) return set_error(HD_ERR_SIZE, "hd_scatter: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    const int N = M * Q;
    const int64_t ecap = (int64_t)CT_SLOT * N + n_ovf;
    char* p = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    double* grad = (double*)p;  p += align256((size_t)N * Z * 2 * 8);
    double* Dm = (double*)p;    p += align256((size_t)G * G * Z * 8);
    double* Y = (double*)p;     p += align256((size_t)G * G * Z * 8);
    double* densT = (double*)p; p += align256((size_t)N * Z * 8);
    int32_t* slot_ovf = (int32_t*)p;  p += align256((size_t)(N + 1) * 4);
    int4* slot_i = (int4*)p;          p += align256((size_t)N * 16);
    double* slot_v = (double*)p;      p += align256((size_t)N * 6 * 8);
    int32_t* pe_q = (int32_t*)p;      p += align256((size_t)ecap * 4);
    double4* pe_geo = (double4*)p;    p += align256((size_t)ecap * 32);
    double* ybuf = (double*)p;        p += align256((size_t)N * Z * 2 * 8);
    double* fdiff = (double*)p;       p += align256((size_t)Z * ecap * 8);
    double* y_hist = (double*)p;
    int32_t* bad = iters_out + Z;   // iters_out has Z + 1 entries

    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, bad, 1);
    hipLaunchKernelGGL(k_ct_slots, dim3(1), dim3(1024), 0, st, N, nbr_ptr, levels_idx, slot_ovf);
    hipLaunchKernelGGL(k_ct_geom, dim3((unsigned)cdiv(N, 256)), dim3(256), 0, st, N, pts, nbr_ptr, nbr_idx, levels_idx,
                       slot_ovf, n_ovf, slot_i, slot_v, pe_q, pe_geo, bad);
    hipLaunchKernelGGL(k_transpose_nz, dim3((unsigned)cdiv((int64_t)N * Z, 256)), dim3(256), 0, st, dens, N, Z, densT);
    hipLaunchKernelGGL(k_ct_fdiff, dim3((unsigned)cdiv((int64_t)N * Z, 256)), dim3(256), 0, st, densT, N, Z, ecap,
                       slot_i, pe_q, fdiff);
    int rc0 = check_launch("k_ct_fdiff");
    if (rc0) return rc0;
    // solution vector (and the level pointers and entry terms) in LDS when they fit
    int dev = 0;
    hipError_t he = hipGetDevice(&dev);
    int max_lds = 0;
    if (he == hipSuccess) he = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (he != hipSuccess) return set_error(HD_ERR_HIP, "hd_scatter: device query: %s", hipGetErrorString(he));
    // the pipelined sweeps when the ring, the step tables and CT_SLOT lanes per vertex of the
    // widest step fit
    const int K = ct_ring(n_levels, sweep_period);
    if (K > 0 && max_step_width > 0) {
        const int nthr_p = (int)(((int64_t)max_step_width * CT_PIPE_LANES + 63) / 64 * 64);
        const int KM = (n_levels + sweep_period - 1) / sweep_period;
        const size_t lds_p = (size_t)N * 16 + (size_t)(nthr_p / CT_PIPE_LANES) * CT_SLOT * 16 +
                             (size_t)(n_levels + 1) * 4 + (size_t)sweep_period * (1 + KM + KM + 1) * 4 +
                             (size_t)sweep_period * max_step_width * 2 + 256;
        if (nthr_p <= CT_MAX_THREADS && max_level_width < 4096 && lds_p <= (size_t)max_lds) {
            if (lds_p > 65536) {
                he = hipFuncSetAttribute((const void*)k_ct_gradients_pipe, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds_p);
                if (he != hipSuccess)
                    return set_error(HD_ERR_HIP, "hd_scatter: LDS attribute: %s", hipGetErrorString(he));
            }
            hipLaunchKernelGGL(k_ct_gradients_pipe, dim3(Z), dim3(nthr_p), lds_p, st, N, Z, slot_i, slot_v, pe_q,
                               pe_geo, fdiff, ecap, levels_ptr, n_levels, sweep_period, K, max_step_width, tol, maxiter,
                               y_hist, grad, iters_out);
            int rcp = check_launch("k_ct_gradients_pipe");
            if (rcp) return rcp;
            goto evaluate;
        }
    }
    {
    // CT_SLOT lanes per vertex of the widest level (the kernel refuses a level wider than its workgroup)
    const int nthr = (int)(((int64_t)max_level_width * CT_SLOT + 63) / 64 * 64);
    if ((uint64_t)ecap * 32 >= (1ull << 32))   // the sweep addresses its arrays with 32-bit byte offsets
        return set_error(HD_ERR_SIZE, "hd_scatter: %d nodes exceed the sweep's 32-bit offsets", N);
    if (nthr > CT_MAX_THREADS)
        return set_error(HD_ERR_SIZE, "hd_scatter: a level of %d vertices exceeds %d threads", max_level_width,
                         CT_MAX_THREADS);
    const size_t aux_bytes = (size_t)nthr * 2 * sizeof(double) + (size_t)(n_levels + 1) * sizeof(int32_t);
    const size_t y_bytes = (size_t)N * 2 * sizeof(double) + aux_bytes;
    if (aux_bytes > (size_t)max_lds) return set_error(HD_ERR_SIZE, "hd_scatter: %d levels exceed the LDS", n_levels);
    const bool lds_y = y_bytes <= (size_t)max_lds;
    const void* kfn = lds_y ? (const void*)k_ct_gradients<true> : (const void*)k_ct_gradients<false>;
    const size_t lds_bytes = lds_y ? y_bytes : aux_bytes;
    if (lds_bytes > 65536) {
        he = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (he != hipSuccess) return set_error(HD_ERR_HIP, "hd_scatter: LDS attribute: %s", hipGetErrorString(he));
    }
    if (lds_y)
        hipLaunchKernelGGL(k_ct_gradients<true>, dim3(Z), dim3(nthr), lds_bytes, st, N, Z, slot_i, slot_v, pe_q,
                           pe_geo, fdiff, ecap, levels_ptr, n_levels, tol, maxiter, ybuf, grad, iters_out);
    else
        hipLaunchKernelGGL(k_ct_gradients<false>, dim3(Z), dim3(nthr), lds_bytes, st, N, Z, slot_i, slot_v, pe_q,
                           pe_geo, fdiff, ecap, levels_ptr, n_levels, tol, maxiter, ybuf, grad, iters_out);
    int rc = check_launch("k_ct_gradients");
    if (rc) return rc;
    }
evaluate:
    dim3 g1((unsigned)cdiv((int64_t)G * G, 256), Z);
    hipLaunchKernelGGL(k_ct_eval, g1, dim3(256), 0, st, dens, N, Z, G, pts, simplices, simplex_g, q_simplex, q_bary,
                       q_nearest, grad, Dm, bad);
    int rc = check_launch("k_ct_eval");
    if (rc) return rc;
    // Y (G x Z*G) = P (G x G) * D (G x Z*G)
    rc = hd_gemm_f64(P, Dm, Y, G, Z * G, G, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_back_interp, dim3((unsigned)cdiv((int64_t)N * Z, 256)), dim3(256), 0, st, Y, N, Z, G, back_idx,
                       back_w, dens);
    return check_launch("k_back_interp");
}

extern "C" int hd_ct_locate(const double* pts, int N, const int32_t* simplices, const double* transform, int S,
                            const double* mgrid, int G, int32_t* q_simplex, double* q_bary, int32_t* q_nearest,
                            void* stream) {
    if (N < 3 || S < 1 || G < 2 || !pts || !simplices || !transform || !mgrid || !q_simplex || !q_bary || !q_nearest)
        return set_error(HD_ERR_ARG, "hd_ct_locate: bad arguments");
    if ((int64_t)G * G > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_ct_locate: G too large");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = (int64_t)G * G;
    hipLaunchKernelGGL(k_ct_fill_i32, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, q_simplex, n, 0x7fffffff);
    hipLaunchKernelGGL(k_ct_raster, dim3((unsigned)cdiv(S, 256)), dim3(256), 0, st, pts, simplices, transform, S, mgrid,
                       G, q_simplex);
    hipLaunchKernelGGL(k_ct_bary_nearest, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, pts, N, transform, mgrid, G,
                       q_simplex, q_bary, q_nearest);
    return check_launch("hd_ct_locate");
}

extern "C" int hd_ct_levels(const int32_t* nbr_ptr, const int32_t* nbr_idx, int N, int32_t* level) {
    if (N < 1 || !nbr_ptr || !nbr_idx || !level) return set_error(HD_ERR_ARG, "hd_ct_levels: bad arguments");
    for (int p = 0; p < N; p++) {   // one more than the highest level of the lower-indexed neighbours
        int32_t l = 0;
        for (int32_t k = nbr_ptr[p]; k < nbr_ptr[p + 1]; k++) {
            const int32_t q = nbr_idx[k];
            if (q >= 0 && q < p && level[q] + 1 > l) l = level[q] + 1;
        }
        level[p] = l;
    }
    return HD_OK;
}
