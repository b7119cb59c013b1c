/*
 * bsg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of one BSGLMM probit Gibbs sweep.
 * This is the checker for the HIP path. Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it. The product never calls
 * it and must fail loudly if its own HIP library is missing.
 *
 * Reference files followed (nicholst/BSGLMM):
 *   layout      bsglmm/main.cu:189-394, 471-502, bsglmm/binCAR.h:12-19
 *   latent      bsglmm/covarGPU.cu:1368-1428 (limits :1414-1418), covar.cpp:302-325
 *   X'Z         bsglmm/covarGPU.cu:327-352
 *   coef update bsglmm/covarGPU.cu:354-392, 124-202; covar.cpp:121-187
 *   pair sums   bsglmm/covarGPU.cu:953-1013, 1075-1087; covar.cpp:220-262
 *   Wishart     bsglmm/covarGPU.cu:1099-1108; cholesky.cpp:13-86, 245-286
 *   KISS        bsglmm/randgen.cpp:9-23, 36-45; main.cu:54, 413
 *
 * The random variates are not the reference's (cuRAND XORWOW cannot be
 * reproduced). Both this oracle and the HIP path draw them from Philox4x32-10
 * with the same (key, counter) scheme, so the uniforms are bit-identical and
 * every conditional update can be compared value by value. Every function
 * that draws variates also has an "injected" form that takes them as arrays;
 * that form is what the tests use to pin this oracle against the reference's
 * own compiled CPU routines.
 */
#ifndef BSG_ORACLE_H
#define BSG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 and the counter scheme -------------------------------- */
enum {
  ORC_TAG_LATENT = 1,
  ORC_TAG_COEF = 2,
  ORC_TAG_GAMMA = 3,
  ORC_TAG_WNORM = 4,
  ORC_TAG_BWM = 5,
  ORC_TAG_FALLBACK = 6
};

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* counter = (idx, sweep, tag | sub << 4, chain) */
void orc_draw(const uint32_t key[2], uint32_t idx, uint32_t sweep, uint32_t tag,
              uint32_t sub, uint32_t chain, uint32_t out[4]);
/* odd 24-bit grid: (2*(x>>9)+1) * 2^-24, exact in float and double */
double orc_u01(uint32_t x);
/* Box-Muller pair from two grid uniforms */
void orc_box_muller(double u0, double u1, double *z0, double *z1);

/* ---- normal tail function and its inverse --------------------------------- */
double orc_Q(double x);             /* P(N(0,1) > x) */
double orc_Qinv(double p);          /* x with Q(x) = p, p in (0,1) */
double orc_logQ(double x);          /* log Q(x), any x */

/* Truncated-normal latent draw by inversion, one uniform:
 * Z ~ N(mu,1) truncated to (eps, 500) if y else (-500, -eps). */
double orc_latent_draw(double mu, int y, double u, double eps);

/* ---- layout (main.cu:189-394, 471-502) ------------------------------------ */
/* counts[0..3] = TOTVOX, TOTVOXp, INDX[0].hostN, INDX[1].hostN.
 * mask is the padded (nrow+2)(ncol+2)(ndepth+2) byte grid. */
void orc_layout_counts(const unsigned char *mask, int nrow, int ncol, int ndepth, int counts[4]);
/* fills hostIdx, hostIdxSC (nsize each), hostVox/hostNBRS per colour, and
 * nbr6[totvox*6]: sorted index (colour-0 list then colour-1 list) of the
 * -x,+x,-y,+y,-z,+z neighbours, -1 when the neighbour is outside the mask. */
void orc_layout_fill(const unsigned char *mask, int nrow, int ncol, int ndepth,
                     int *host_idx, int *host_idx_sc, int *vox0, int *vox1,
                     unsigned char *nbrs0, unsigned char *nbrs1, int *nbr6);

/* ---- one sweep, reference layout ----------------------------------------- */
/* Problem in the reference's own layout:
 *   covar[P*N]       covariate-major, covar[c*N+i]   (read_data.cpp:299-302)
 *   data[N*TOTVOX]   y[i*TOTVOX + hostIdx[vox]]      (read_data.cpp:385)
 *   spatCoef[P*TOTVOXp] beta[c*TOTVOXp + hostIdxSC[vox]]   (mcmc.cpp:227)
 *   WM[TOTVOX]       1.0 when no WM image             (read_data.cpp:531-534)
 */
typedef struct {
  int nrow, ncol, ndepth;
  int nsubs, ncovar, totvox, totvoxp;
  int n_col[2];
  const int *host_idx, *host_idx_sc;
  const int *vox[2];
  const unsigned char *nbrs[2];
  const float *covar;
  const float *xxprime;       /* P*P */
  const unsigned char *data;
  const float *wm;
  uint32_t key[2];
  uint32_t chain;
} orc_problem;

/* latent update; Z_out[i*TOTVOX + hostIdx] (double); b = WM coefficient.
 * method 0: every draw by inversion from uniform l of the quad's block.
 * method 1: Box-Muller candidates from the quad's block, (u0,u1) and (u2,u3);
 * a candidate outside the interval of y is redrawn by inversion from
 * block (FALLBACK, sub = l). Both are exact truncated-normal samplers. */
void orc_update_z(const orc_problem *pb, const float *spat_coef, float b, double eps,
                  uint32_t sweep, int method, double *z_out);
/* M[c*TOTVOX + hostIdx] = sum_i covar[c,i] (Z[i,idx] - b WM[idx]) */
void orc_xtz(const orc_problem *pb, const float *z, float b, double *m_out);
/* coefficient update of one colour, reference math (chol, invert, chol).
 * spat_coef (double, P*TOTVOXp) is updated in place for that colour.
 * m is X'(Z-bWM) as above. If z_inj is non-NULL it holds P normals per list
 * slot (z_inj[slot*P + c]) and replaces the Philox normals. */
void orc_update_coef(const orc_problem *pb, int colour, const double *m, const float *prec,
                     double *spat_coef, uint32_t sweep, const double *z_inj);
/* neighbour-pair scale sums T (P*P, symmetric), Spat_Coef_Prec semantics */
void orc_pair_sums(const orc_problem *pb, const float *spat_coef, double *t_out);
/* Wishart draw of the CAR precision from T: S = (T+I)^-1, Lambda ~ W(df, S).
 * chi2_inj/norm_inj (optional) replace the Philox variates: chi2_inj[i] is
 * chi-square(df-i); norm_inj[i*P+j] (j<i) the below-diagonal normals. */
void orc_wishart(int p, const double *t, int df, const uint32_t key[2], uint32_t sweep,
                 uint32_t chain, const double *chi2_inj, const double *norm_inj,
                 double *lambda_out);
/* WM coefficient (covarGPU.cu:588-724, GPU semantics: flat prior):
 *   S = sum_j WM_j sum_i (Z_ij - x_i'b_j),  W = N sum_j WM_j^2  (z as orc_update_z) */
void orc_wm_stat(const orc_problem *pb, const float *spat_coef, const double *z, double *S, double *W);
/* b = S/W + sqrt(1/W) n, n from Philox (0, sweep, ORC_TAG_BWM, chain) by Box-Muller */
double orc_wm_draw(double S, double W, const uint32_t key[2], uint32_t sweep, uint32_t chain);
/* The reference CPU path's WM-coefficient statistic (updateBeta,
 * mcmc.cpp:1099-1140, which cannot be compiled without CUDA): raster order
 * over the padded grid, float Kahan sums, z float[N][TOTVOX], wm float[TOTVOX].
 * Returns the posterior mean and variance of b before its normal draw
 * (var = 1 / (N sum WM^2 + prior_prec), mean = var (S + prior_mean prior_prec)). */
void orc_cpu_wm_stat(const orc_problem *pb, const float *spat_coef, const float *z, const float *wm,
                     float prior_mean, float prior_prec, float *mean, float *var);
/* Marsaglia-Tsang gamma(a,1), a >= 1, from Philox stream (row, sweep, chain) */
double orc_gamma(double a, const uint32_t key[2], uint32_t row, uint32_t sweep, uint32_t chain);

/* ---- dense helpers (cholesky.cpp:13-86 semantics, row-major lower) -------- */
int orc_chol(double *a, int n);
void orc_chol_invert(int n, double *h);

/* ---- host KISS (randgen.cpp:9-23 with M = exp(-64 ln 2), main.cu:54) ----- */
double orc_kiss(unsigned long *seed);
unsigned long orc_runiform_long_n(unsigned long n, unsigned long *seed);

#ifdef __cplusplus
}
#endif
#endif
