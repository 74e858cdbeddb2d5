/*
 * holohip.h -- C-ABI of libholohip.so, the MI355X (gfx950) implementation of
 * holodeck's semi-analytic GW-background hot path.
 *
 * Conventions (every entry point):
 *   - pointers are DEVICE pointers to contiguous row-major (C-order) arrays;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream); every
 *     call is asynchronous on that stream;
 *   - the return value is a status: 0 = OK, <0 = error; hd_last_error() gives
 *     a message for the calling thread;
 *   - no entry point allocates memory visible to the caller; entry points that
 *     need scratch take a caller-provided workspace whose size is reported by
 *     the matching *_workspace_size() query (hd_realize_noalloc instead uses a
 *     fixed device array of the library);
 *   - FP64 in and out, as the reference's typed memoryviews (double[:,...]).
 *
 * Each declaration cites the reference interface it replaces (holodeck v1.6,
 * paths relative to the reference repository root).
 */
#ifndef HOLOHIP_H
#define HOLOHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message for the last error on the calling thread ("" if none). */
const char *hd_last_error(void);
/* Library ABI version. */
int hd_version(void);

/* Launch timing (no counterpart in the reference; for benchmarks).  While
 * enabled, selected launch sequences record a HIP event pair on their stream:
 * "hd_realize" (the whole hd_realize sequence) and "k_gz_draw" (its draw
 * kernel).  hd_timing_enable clears earlier records; hd_timing_read waits for
 * the recorded events and returns the summed milliseconds and the launch
 * count for one name. */
int hd_timing_enable(int on);
int hd_timing_read(const char *name, double *total_ms, int *launches);

/* ------------------------------------------------------------------------
 * Cosmology (flat LambdaCDM, no radiation).  Replaces the cosmopy/astropy
 * `holo.cosmo` instance (holodeck/__init__.py:81-86): the interpolation grids
 * `_grid_z/_grid_age/_grid_dcom` consumed by sam_cyutils.pyx:476,494 and the
 * exact `age`, `tage_to_z`, `dtdz`, `comoving_distance` used on the path
 * (sam.py:347, utils.py:1772-1808, gravwaves.py:713-714).
 * ---------------------------------------------------------------------- */
/* Comoving-distance Hermite table: dc_tab, dc_der each (n_tab+1) doubles,
 * in units of c/H0, uniform in u = (1+z)^-1/2 on [0,1]. */
int hd_cosmo_build_dcom_table(double om0, int n_tab, double *dc_tab, double *dc_der, void *stream);
/* age [s] and comoving distance [cm] at the n redshifts z (table nodes). */
int hd_cosmo_eval(double om0, double h, const double *dc_tab, const double *dc_der, int n_tab, const double *z,
                  int64_t n, double *age, double *dcom, double *dtdz, void *stream);

/* ------------------------------------------------------------------------
 * Static binary density  <- Semi_Analytic_Model.static_binary_density
 * (holodeck/sams/sam.py:280-398) with GSMF_Schechter (components.py:110-172),
 * GPF_Power_Law (520-583) / GMT_Power_Law (629-675) or GMR_Illustris
 * (380-482), MMBulge_Standard + BF_Constant (host_relations.py:166-195,
 * 624-771) and GMT redshift delay utils.redz_after (utils.py:1772-1808).
 * Writes dens (M,Q,Z) [Mpc^-3] BEFORE scatter and before zeroing stalled
 * bins, plus zprime (M,Q,Z) and gmt_time (M,Q,Z) [s].
 * `par` is a HOST array of HD_DENS_NPAR doubles (layout:
 * holodeck_amd/csrc/density_par.h); dc_tab/dc_der from
 * hd_cosmo_build_dcom_table.
 * ---------------------------------------------------------------------- */
#define HD_DENS_NPAR 48
int hd_static_density(const double *mtot, int M, const double *mrat, int Q, const double *redz, int Z,
                      const double *par, const double *dc_tab, const double *dc_der, int n_tab,
                      double *dens, double *zprime, double *gmt_time, void *stream);
/* dens[zprime < 0] = 0  (sam.py:392-394) */
/* <- Semi_Analytic_Model.mass_stellar (holodeck/sams/sam.py:250-278): stellar masses of the
 * primary, their ratio and total per (M,Q,Z) node (redshift-independent, broadcast over Z) from
 * the M-Mbulge parameters in `par` (the host array of hd_static_density). */
int hd_mass_stellar(const double *mtot, int M, const double *mrat, int Q, int Z, const double *par, double *mstar_pri,
                    double *mstar_rat, double *mstar_tot, void *stream);
int hd_zero_stalled(double *dens, const double *zprime, int64_t n, void *stream);

/* ------------------------------------------------------------------------
 * M-Mbulge scatter  <- add_scatter_to_masses (sam.py:1291-1394) with
 * utils._get_rolled_weights/_scatter_with_weights (utils.py:416-488).
 * dens (M,Q,Z) is updated in place, all Z slices at once.  The plan arrays
 * are geometry built once per grid shape on the host (holodeck_amd/scatter.py):
 *   pts (N,2) node coordinates (log10 m1, log10 m2), N = M*Q, C order;
 *   nbr_ptr (N+1) / nbr_idx: vertex adjacency of the Delaunay triangulation;
 *   levels_ptr (n_levels+1) / levels_idx (N): Gauss-Seidel level schedule;
 *   max_level_width: the most vertices of one level (at most 128);
 *   n_ovf: adjacency entries beyond the eighth of each vertex, summed;
 *   sweep_period: one more than the largest level difference across an edge
 *     (0: the Gauss-Seidel sweeps run one after another); with it sweep s + 1
 *     starts sweep_period levels behind sweep s (pipelined sweeps, the same
 *     arithmetic), and max_step_width is the most vertices such a step holds
 *     (the levels l, l + P, l + 2P, ... together);
 *   simplices (S,3), simplex_g (S,3) cross-boundary derivative directions;
 *   q_simplex (G*G) (-1 outside the hull), q_bary (G*G,3), q_nearest (G*G);
 *   P (G,G) = (W W)^T for the Gaussian-CDF weight matrix W (see hd_gemm_f64);
 *   back_idx (N,4) regular-grid points, back_w (N,4) bilinear weights.
 * iters_out: Z+1 int32 -- Gauss-Seidel sweeps per slice (0: no convergence
 * within maxiter) and, last, the count of values still NaN/negative after the
 * nearest-node fill (the reference raises in that case, sam.py:1377-1383),
 * with bit 20 set when the adjacency has more than n_ovf such entries and
 * bit 21 when a level has more than max_level_width vertices (nothing
 * computed then).
 * ---------------------------------------------------------------------- */
size_t hd_scatter_workspace_size(int M, int Q, int Z, int G, int n_ovf, int n_levels, int sweep_period);
int hd_scatter(double *dens, int M, int Q, int Z, int G, const double *pts, const int32_t *nbr_ptr,
               const int32_t *nbr_idx, const int32_t *levels_ptr, const int32_t *levels_idx, int n_levels,
               int max_level_width, int n_ovf, int sweep_period, int max_step_width,
               const int32_t *simplices, const double *simplex_g, const int32_t *q_simplex, const double *q_bary,
               const int32_t *q_nearest, const double *P, const int32_t *back_idx, const double *back_w, double tol,
               int maxiter, void *workspace, size_t ws_bytes, int32_t *iters_out, void *stream);
/* The scatter geometry's device part, from the triangulation of the N nodes (pts (N,2);
 * simplices (S,3) and their affine transforms (S,3,2), as scipy.spatial.Delaunay.transform):
 * for every point xi[i G + j] = (mgrid[i], mgrid[j]) of the regular grid, the simplex holding
 * it (-1 outside the hull; Qhull's inside test with eps = 100 DBL_EPSILON, the lowest simplex
 * for a point on a shared edge) with its barycentric coordinates q_bary (G*G,3), and the
 * nearest node q_nearest (squared distance, lowest index on a tie) -- the lookups of
 * CloughTocher2DInterpolator and NearestNDInterpolator in add_scatter_to_masses
 * (sam.py:1365-1376). */
int hd_ct_locate(const double *pts, int N, const int32_t *simplices, const double *transform, int S,
                 const double *mgrid, int G, int32_t *q_simplex, double *q_bary, int32_t *q_nearest, void *stream);
/* Host: the Gauss-Seidel level of every node (one more than the highest level of its
 * lower-indexed neighbours, 0 without any), from the vertex adjacency (nbr_ptr (N+1), nbr_idx). */
int hd_ct_levels(const int32_t *nbr_ptr, const int32_t *nbr_idx, int N, int32_t *level);
/* C (m,n) = A (m,k) B (k,n), row-major FP64 (the scatter's convolution GEMM;
 * the reference evaluates it with numpy.einsum, utils.py:455-461). */
int hd_gemm_f64(const double *A, const double *B, double *C, int m, int n, int k, void *stream);
/* <- the weight matrix of utils._get_rolled_weights (holodeck/utils.py:464-488), transposed:
 * wT (G,G) with wT[k,j] = dm[k - j + G - 1] for the 2G-1 Gaussian-CDF differences dm. */
int hd_scatter_weights_T(const double *dm, int G, double *wT, void *stream);

/* ------------------------------------------------------------------------
 * Binary evolution  <- holodeck/sams/sam_cyutils.pyx
 * ---------------------------------------------------------------------- */
/* find_2pwl_hardening_norm (sam_cyutils.pyx:289-353): brentq per (mtot,mrat).
 * norm (may be NULL) receives 10**norm_log10, as Fixed_Time_2PL_SAM stores
 * it (hardening.py:1413-1416). */
int hd_find_2pwl_norm(double target_time, const double *mtot, const double *mrat, int64_t n, double sepa_init,
                      double rchar, double gamma_inner, double gamma_outer, int nsteps, double *norm_log10,
                      double *norm, void *stream);
/* integrate_binary_evolution_2pwl (sam_cyutils.pyx:401-418), elementwise over n: the 300-step
 * (nsteps) lifetime [s] from sepa_init to 3 Schwarzschild radii at normalisation 10**norm_log10. */
int hd_lifetime_2pwl(const double *norm_log10, const double *mtot, const double *mrat, int64_t n, double sepa_init,
                     double rchar, double gamma_inner, double gamma_outer, int nsteps, double *time, void *stream);
/* hard_func_2pwl_gw (sam_cyutils.pyx:271-286), elementwise over n. */
int hd_hard_func_2pwl_gw(const double *mtot, const double *mrat, const double *sepa, const double *norm, int64_t n,
                         double rchar, double gamma_inner, double gamma_outer, double *dadt, void *stream);
/* _dynamic_binary_number_at_fobs_gw (sam_cyutils.pyx:784-899). */
int hd_dbn_gw(const double *fobs_orb, int F, const double *nden, const double *mtot, int M, const double *mrat, int Q,
              int Z, const double *redz_prime, const double *grid_z, const double *grid_dcom, int n_interp,
              double *redz_final, double *diff_num, void *stream);
/* _dynamic_binary_number_at_fobs_2pwl (sam_cyutils.pyx:506-781).
 * fobs_sorted != 0 promises fobs_orb is non-decreasing (enables a binary search). */
int hd_dbn_2pwl(const double *fobs_orb, int F, int fobs_sorted, double sepa_init, int num_steps,
                const double *hard_norm, double rchar, double gamma_inner, double gamma_outer, const double *nden,
                const double *mtot, int M, const double *mrat, int Q, const double *redz, int Z,
                const double *gmt_time, const double *grid_z, const double *grid_dcom, const double *grid_age,
                int n_interp, double *redz_final, double *diff_num, void *stream);
/* _integrate_differential_number_3dx1d (sam_cyutils.pyx:166-216):
 * log10_mtot (M), mrat (Q), redz (Z), dln_freq (F), dnum (M,Q,Z,F) -> numb (M-1,Q-1,Z-1,F). */
int hd_integrate_3dx1d(const double *log10_mtot, int M, const double *mrat, int Q, const double *redz, int Z,
                       const double *dln_freq, int F, const double *dnum, double *numb, void *stream);

/* ------------------------------------------------------------------------
 * Strain  <- gravwaves.char_strain_sq_from_bin_edges_redz (gravwaves.py:694-725)
 * mt_cent (M-1), mr_cent (Q-1) bin centres; fc (F) orbital bin centres,
 * fc_over_df (F); redz_final (M,Q,Z,F) -> h2fdf (M-1,Q-1,Z-1,F).
 * gw_src_const = 8 G^(5/3) pi^(2/3) / sqrt(10) / c^4 (utils.py:40, astropy G, c).
 * ---------------------------------------------------------------------- */
int hd_h2fdf_redz(const double *mt_cent, int M, const double *mr_cent, int Q, int Z, const double *fc,
                  const double *fc_over_df, int F, const double *redz_final, double gw_src_const, double om0, double h,
                  const double *dc_tab, const double *dc_der, int n_tab, double *h2fdf, void *stream);

/* Per-bin parameters for the loudest walk with parameters
 * (single_sources.py:104-139): out (7, B, F) = [mt, mr, rz (bin centres),
 * redz_final (midpointed; -1 where <= 0), dcom_final [cm] (+inf where
 * redz_final <= 0), sepa_final [cm], angs_final [rad]]; fc (F) orbital centres. */
int hd_bin_params_redz(const double *mt_cent, int M, const double *mr_cent, int Q, const double *rz_cent, int Z,
                       const double *fc, int F, const double *redz_final, double om0, double h,
                       const double *dc_tab, const double *dc_der, int n_tab, double *out, void *stream);

/* <- gravwaves.gwb_ideal (holodeck/gravwaves.py:619-663) for Semi_Analytic_Model.gwb_ideal
 * (holodeck/sams/sam.py:837-870): the idealised GWB h_c(f) of a continuous, GW-driven
 * population.  ndens (M,Q,Z) [Mpc^-3] at the grid nodes, log10 and linear total-mass edges [g],
 * mass-ratio edges, redshifts: (Z) edges, or with redz_3d != 0 a per-node z' (M,Q,Z) (the
 * reference's redz_prime); constants G, c [cgs] and Mpc [cm] as the caller's module defines
 * them.  sum != 0: out (F) h_c; workspace of hd_gwb_ideal_workspace_size() bytes.
 * sum == 0: out (M-1,Q-1,Z-1,F) per-cell h_c, no workspace. */
size_t hd_gwb_ideal_workspace_size(void);
int hd_gwb_ideal(const double *ndens, const double *log10_mtot, const double *mtot, int M, const double *mrat, int Q,
                 const double *redz, int redz_3d, int Z, const double *fobs_gw, int F, double nwtg, double splc,
                 double mpc, int sum, double *out, void *workspace, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------------
 * Ordering  <- single_sources.ss_gws_redz's argsort(-h2fdf[..., 0])
 * (single_sources.py:89-93), made stable: descending by key, ties by
 * ascending bin index.  key (B, stride) column `col`; order (B) int64 out.
 * ---------------------------------------------------------------------- */
size_t hd_order_workspace_size(int64_t B);
int hd_order_desc(const double *key, int64_t B, int64_t stride, int64_t col, int64_t *order, void *workspace,
                  size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------------
 * Realizations (holodeck/cyutils.pyx).  number, h2fdf are (B,F) with B the
 * flat C-order bin index over (M-1,Q-1,Z-1).
 *
 * Loud-source selection modes:
 *   HD_SEL_SORTED : walk bins in `order` (loudest first); the first L source
 *                   instances (with multiplicity) are the loud sources; the
 *                   rest is background  <- _loudest_hc_from_sorted
 *                   (cyutils.pyx:1266-1344); with HD_FLAG_SKIP_ZERO_H also
 *                   _loudest_hc_and_par_from_sorted_redz (1615-1767).
 *   HD_SEL_ARGMAX : per frequency, the L largest h2fdf instances among bins
 *                   with n>0, ties to the lower bin index; L=1 is
 *                   _ss_bg_hc(_and_par) (cyutils.pyx:935-1178), where the
 *                   background is sum(n*h) - max.
 *   HD_SEL_NONE   : no loud sources; hc2bg = sum(n*h)  <- _sam_poisson_gwb
 *                   (cyutils.pyx:862-897).
 * Outputs: hc2ss (F,R,L) and ssidx (F,R,L) (flat bin index, -1 if empty),
 * hc2bg (F,R); bgpar (npar,F,R) strain-weighted background means of
 * par (npar,B,F) when npar>0.
 * ---------------------------------------------------------------------- */
#define HD_SEL_SORTED 0
#define HD_SEL_ARGMAX 1
#define HD_SEL_NONE 2
#define HD_FLAG_SKIP_ZERO_H 1

/* Injected counts (B,F,R) (deterministic; the parity path). */
size_t hd_realize_counts_workspace_size(int64_t B, int F, int R, int L, int npar);
int hd_realize_counts(const double *counts, const double *h2fdf, const int64_t *order, int64_t B, int F, int R, int L,
                      int mode, int flags, const double *par, int npar, double *hc2ss, int64_t *ssidx, double *hc2bg,
                      double *bgpar, void *workspace, size_t ws_bytes, void *stream);

/* On-GPU Poisson realizations (counter-based RNG keyed by seed and
 * (bin, frequency, realization)); normal_threshold as the reference's 1e10. */
size_t hd_realize_workspace_size(int64_t B, int F, int R, int L, int npar);
int hd_realize(const double *number, const double *h2fdf, const int64_t *order, int64_t B, int F, int R, int L,
               int mode, int flags, double normal_threshold, uint64_t seed, const double *par, int npar,
               double *hc2ss, int64_t *ssidx, double *hc2bg, double *bgpar, void *workspace, size_t ws_bytes,
               void *stream);

/* Test hook of hd_realize (no counterpart in the reference, whose Cython loops draw each count
 * and use it at once, cyutils.pyx:1266-1344, 1615-1767, 1068-1178, 862-897): the same launch
 * sequence and outputs as hd_realize for the same seed, and in addition every count the draw
 * kernels drew, counts_out (B,F,R) (0 for bins that are never drawn: number 0, or excluded by the
 * mode's rules), so that the reference's walks can be run on exactly those counts. */
int hd_realize_record_counts(const double *number, const double *h2fdf, const int64_t *order, int64_t B, int F,
                             int R, int L, int mode, int flags, double normal_threshold, uint64_t seed,
                             const double *par, int npar, double *hc2ss, int64_t *ssidx, double *hc2bg,
                             double *bgpar, double *counts_out, void *workspace, size_t ws_bytes, void *stream);

/* hd_realize without a workspace, for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  It has no counterpart in
 * the reference.  Same arguments as hd_realize minus the workspace; same launch sequence, random
 * streams and outputs for the same seed.  Its scratch is a fixed device array of
 * HD_NOALLOC_SCRATCH_BYTES on each GPU the library runs on, held until the process ends, so it
 * takes only sizes with hd_realize_workspace_size(B, F, R, L, npar) <= HD_NOALLOC_SCRATCH_BYTES
 * (without parameters about 2.7 M bin-frequency cells at R = 500, L = 10 and 3.4 M at R = 100,
 * L = 5, e.g. 120,000 bins x 20 frequencies x 100 realizations);
 * larger ones return HD_ERR_SIZE.  Every call on one GPU uses that GPU's one array: calls that
 * overlap in time on one GPU (other streams, other threads) race, and the caller must
 * serialise them.  Allocates nothing (no hipMalloc, hipFree or allocating library call).  The
 * stream must belong to the current device. */
#define HD_NOALLOC_SCRATCH_BYTES (256ull * 1024ull * 1024ull)
int hd_realize_noalloc(const double *number, const double *h2fdf, const int64_t *order, int64_t B, int F, int R,
                       int L, int mode, int flags, double normal_threshold, uint64_t seed, const double *par, int npar,
                       double *hc2ss, int64_t *ssidx, double *hc2bg, double *bgpar, void *stream);

/* Philox4x32-10 block function (host), for known-answer tests of the
 * generator the realization kernels use. */
void hd_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
/* xoshiro128** (host): advances state[4] n times, writing the outputs to out (may be NULL).
 * The realization kernels draw from one such stream per (tile, frequency, realization), seeded
 * by Philox; exported for tests of the generator. */
void hd_xoshiro128ss(uint32_t *state, int64_t n, uint32_t *out);
/* Test hooks of the realization kernels' sampler (no counterpart in the reference, whose
 * numpy Poisson draws they replace): the CDF table of Poisson(lam) built by one wave
 * (kb_kt_n = {kb, kt, n}: the support starts at kb, the table holds the counts kt .. kt + n - 1,
 * cdf[j] = floor(CDF(kt + j) 2^64) with the CDF summed from kb, saturating; cap at most 8000),
 * and draws from such a table with given words, tail walks included (tab = the n high words then
 * the n low words, copied to LDS as the kernels hold it; lam, kb, kt those of the table; two
 * words per draw; count[i] the count drawn, used[i] the words it took). */
int hd_poisson_cdf_table(double lam, int cap, int *kb_kt_n, uint64_t *cdf, void *stream);
/* The same table on the host (no GPU work): support start kb (dropping less than 2^-56 below),
 * table [kt, kt + n) covering lam -/+ 4 sqrt(lam) and the CDF below it; n = 0 when it needs more
 * than cap entries or the support's end would not fit an int. */
int hd_poisson_window(double lam, int cap, int *kb, int *kt, int *n);
int hd_table_draw_words(const uint32_t *tab, int n, double lam, int kb, int kt, const uint32_t *words, int64_t m,
                        int *count, int *used, void *stream);
/* The same for two tables drawn together (the kernels' interleaved pair): four words per pair,
 * taken as x_a, x_b, then w_a and w_b where a tie needs them, then where a tail walk needs them;
 * count[2i], count[2i+1] the counts drawn. */
int hd_table_draw_pair_words(const uint32_t *tab, int n, double lam, int kb, int kt, const uint32_t *tab2, int n2,
                             double lam2, int kb2, int kt2, const uint32_t *words, int64_t m, int *count, int *used,
                             void *stream);

/* ------------------------------------------------------------------------
 * Learned surrogate (MLP: parameters -> log10 h_c percentiles).  No
 * counterpart in the reference, whose emulator is the Gaussian process of
 * holodeck/gps/gp_utils.py; the surrogate sits beside it.  BF16 operands
 * (passed as void*: 2-byte bfloat16), FP32 accumulation and master weights.
 * ---------------------------------------------------------------------- */
#define HD_MLP_EPI_STORE 0     /* out_f32 = acc (+ bias)                          */
#define HD_MLP_EPI_BIAS_SILU 1 /* z = acc + bias -> out_f32 (opt.); silu(z) -> out_bf16 */
#define HD_MLP_EPI_BIAS_BF16 2 /* acc (+ bias) -> out_bf16 (and out_f32 if given)  */
#define HD_MLP_EPI_SILU_BWD 3  /* acc * silu'(zin) -> out_bf16                     */
/* C (M x N) = A (M x K) . B^T, B given as (N x K); element (r, k) of A at
 * A[r*sam + k*sak], of B at B[r*sbn + k*sbk]; C row stride ldo. */
int hd_mlp_gemm_bf16(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbn, int64_t sbk, int M, int N,
                     int K, int epi, const float *bias, const float *zin, float *out_f32, void *out_bf16,
                     int64_t ldo, void *stream);
/* out[n] = sum_m a[m, n] (bias gradients) */
int hd_mlp_colsum_bf16(const void *a, int M, int N, float *out, void *stream);
/* dout = 2 (y - t) / n (bf16); *loss += mean squared error (FP32, accumulated) */
int hd_mlp_mse(const float *y, const float *t, int64_t n, void *dout, float *loss, void *stream);
/* Adam step `step` (>= 1) on FP32 master weights w with gradient g and moments
 * m, v; writes the bf16 copy w_bf16 when not NULL */
int hd_mlp_adam(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                int step, void *w_bf16, void *stream);
int hd_f32_to_bf16(const float *a, int64_t n, void *b, void *stream);

/* ------------------------------------------------------------------------
 * Eccentric SAM GWB over one eccentricity evolution e(a) shared by all binaries.
 * Inputs as the reference's Cython signatures: ndens (M,Q,Z) at the grid nodes
 * [Mpc^-3], log10 of the total-mass edges [g], mass-ratio and redshift edges, comoving
 * distances (Z) [Mpc], observer GW frequencies (F) [1/s], the evolution's separations
 * (E) [cm] (decreasing) and eccentricities (E).
 * ---------------------------------------------------------------------- */
size_t hd_eccen_workspace_size(int M, int Z, int E);
/* <- cyutils.sam_calc_gwb_single_eccen (holodeck/cyutils.pyx:361-597): gwb (F, nharms),
 * the characteristic strain squared per frequency and harmonic n = 1..nharms. */
int hd_gwb_eccen(const double *ndens, const double *mtot_log10, int M, const double *mrat, int Q, const double *redz,
                 const double *dcom, int Z, const double *gwfobs, int F, const double *sepa_evo,
                 const double *eccen_evo, int E, int nharms, double *gwb, void *workspace, size_t ws_bytes,
                 void *stream);
/* <- the per-(node, frequency, harmonic) terms of cyutils.sam_calc_gwb_single_eccen_discrete
 * (holodeck/cyutils.pyx:766-838) for harmonics h0+1 .. h0+hc: number (M*Q*Z, F*hc) is the
 * Poisson mean `number_term` and hval (M*Q*Z, F*hc) the strain factor `hterm / weight` of each
 * drawn count; column aa*hc + (n-1-h0).  Zero where the reference draws nothing or adds only
 * zeros.  Requires M*Z <= 65535.  The realizations are hd_realize with HD_SEL_NONE over these
 * columns. */
int hd_eccen_number(const double *ndens, const double *mtot_log10, int M, const double *mrat, int Q,
                    const double *redz, const double *dcom, int Z, const double *gwfobs, int F, const double *sepa_evo,
                    const double *eccen_evo, int E, int h0, int hc, double *number, double *hval, void *workspace,
                    size_t ws_bytes, void *stream);
/* <- the Bessel pair J_{n-2}(ne), J_{n-1}(ne) that cyutils.pyx:181-182 takes from
 * scipy.special.jv: out[2i], out[2i+1] = J_lo[i](x[i]), J_{lo[i]+1}(x[i]) for -1 <= lo <= 100000,
 * 0 < x <= 1e5 (NaN outside).  nharms and h0+hc above are limited to 100000 as well. */
int hd_bessel_jpair(const int *lo, const double *x, int64_t n, double *out, void *stream);
/* <- cyutils.gw_freq_dist_func__scalar_scalar (holodeck/cyutils.pyx:145-199): g(n, e) elementwise. */
int hd_gw_freq_dist(const int *nn, const double *ee, int64_t n, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* HOLOHIP_H */
