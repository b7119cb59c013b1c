/*
 * bsg_oracle.c -- TEST INFRASTRUCTURE ONLY (see bsg_oracle.h).
 *
 * FP64 restatement of the BSGLMM probit Gibbs sweep. Each function names the
 * reference lines it follows. Nothing here is linked into the product.
 */
#include "bsg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                         */
/* ------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void orc_draw(const uint32_t key[2], uint32_t idx, uint32_t sweep, uint32_t tag,
              uint32_t sub, uint32_t chain, uint32_t out[4])
{
  uint32_t ctr[4] = {idx, sweep, tag | (sub << 4), chain};
  orc_philox4x32_10(ctr, key, out);
}

double orc_u01(uint32_t x)
{
  return (double)((x >> 9) * 2u + 1u) * (1.0 / 16777216.0);
}

void orc_box_muller(double u0, double u1, double *z0, double *z1)
{
  double r = sqrt(-2.0 * log(u0));
  double th = 2.0 * M_PI * u1;
  *z0 = r * cos(th);
  *z1 = r * sin(th);
}

/* ------------------------------------------------------------------------- */
/* Normal tail Q(x) = P(N(0,1) > x), log Q and the inverse.                   */
/* ------------------------------------------------------------------------- */
double orc_Q(double x)
{
  return 0.5 * erfc(x * M_SQRT1_2);
}

double orc_logQ(double x)
{
  if (x < 30.0)
    return log(orc_Q(x));
  /* asymptotic series Q(x) = phi(x)/x (1 - 1/x^2 + 3/x^4 - 15/x^6 + 105/x^8 - 945/x^10) */
  double r = 1.0 / (x * x);
  double s = 1.0 + r * (-1.0 + r * (3.0 + r * (-15.0 + r * (105.0 + r * -945.0))));
  return -0.5 * x * x - log(x) - 0.5 * log(2.0 * M_PI) + log(s);
}

/* Solve log Q(x) = lp for lp <= log(0.5): rational start (Abramowitz & Stegun
 * 26.2.23, |error| < 4.5e-4) then Newton on log Q, which keeps full relative
 * accuracy deep in the tail. */
static double qinv_log(double lp)
{
  double t = sqrt(-2.0 * lp);
  double x = t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                     (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308)));
  for (int it = 0; it < 12; it++) {
    double lq = orc_logQ(x);
    double lphi = -0.5 * x * x - 0.5 * log(2.0 * M_PI);
    double dx = (lq - lp) * exp(lq - lphi); /* g/(-g') with g = logQ - lp */
    x += dx;
    if (fabs(dx) < 1e-16 * (1.0 + fabs(x)))
      break;
  }
  return x;
}

double orc_Qinv(double p)
{
  if (p > 0.5)
    return -qinv_log(log1p(-p));
  return qinv_log(log(p));
}

/* Z ~ N(mu,1) truncated to (eps,500) when y=1 and (-500,-eps) when y=0
 * (covarGPU.cu:1414-1422; covar.cpp:316-321 uses eps = 0).
 * With s = +1 (y=1) or -1 (y=0), Z = mu + s W and W ~ N(0,1) truncated to
 * (a, b), a = eps - s mu, b = 500 - s mu; W = Qinv(Q(b) + u (Q(a) - Q(b))). */
double orc_latent_draw(double mu, int y, double u, double eps)
{
  double s = y ? 1.0 : -1.0;
  double a = eps - s * mu;
  double b = 500.0 - s * mu;
  double w;
  if (a >= 0.0) {
    double lqa = orc_logQ(a), lqb = orc_logQ(b);
    double lp = lqa + log(u + (1.0 - u) * exp(lqb - lqa));
    w = qinv_log(lp);
  } else {
    double qt = orc_Q(-a);          /* 1 - Q(a) */
    double qb = orc_Q(b), qnb = orc_Q(-b);
    double p = qb + u * (qnb - qt); /* Q(b) + u (Q(a) - Q(b)) */
    if (p <= 0.5)
      w = qinv_log(log(p));
    else
      w = -qinv_log(log((1.0 - u) * qnb + u * qt)); /* 1 - p, accurately */
  }
  if (w < a)
    w = a;
  if (w > b)
    w = b;
  return mu + s * w;
}

/* ------------------------------------------------------------------------- */
/* Layout (main.cu:189-394 and cnt_nbrs main.cu:471-502).                    */
/* ------------------------------------------------------------------------- */
static int colour_of(int i, int j, int k)
{
  /* main.cu:221-247: colour 0 iff i+j+k is odd (padded, 1-based coordinates) */
  return ((i + j + k) & 1) ? 0 : 1;
}

void orc_layout_counts(const unsigned char *mask, int nrow, int ncol, int ndepth, int counts[4])
{
  int sx = nrow + 2, sxy = (nrow + 2) * (ncol + 2);
  int nsize = sxy * (ndepth + 2);
  unsigned char *mskp = (unsigned char *)calloc((size_t)nsize, 1);
  int totvox = 0, n[2] = {0, 0};
  for (int k = 1; k < ndepth + 1; k++)
    for (int j = 1; j < ncol + 1; j++)
      for (int i = 1; i < nrow + 1; i++) {
        int id = i + sx * j + sxy * k;
        if (mask[id]) {
          totvox++;
          n[colour_of(i, j, k)]++;
          mskp[id] = mskp[id - 1] = mskp[id + 1] = 1;
          mskp[id - sx] = mskp[id + sx] = mskp[id - sxy] = mskp[id + sxy] = 1;
        }
      }
  int totvoxp = 0;
  for (int id = 0; id < nsize; id++)
    totvoxp += mskp[id] != 0;
  free(mskp);
  counts[0] = totvox;
  counts[1] = totvoxp;
  counts[2] = n[0];
  counts[3] = n[1];
}

void orc_layout_fill(const unsigned char *mask, int nrow, int ncol, int ndepth,
                     int *host_idx, int *host_idx_sc, int *vox0, int *vox1,
                     unsigned char *nbrs0, unsigned char *nbrs1, int *nbr6)
{
  int sx = nrow + 2, sxy = (nrow + 2) * (ncol + 2);
  int nsize = sxy * (ndepth + 2);
  unsigned char *mskp = (unsigned char *)calloc((size_t)nsize, 1);
  for (int k = 1; k < ndepth + 1; k++)
    for (int j = 1; j < ncol + 1; j++)
      for (int i = 1; i < nrow + 1; i++) {
        int id = i + sx * j + sxy * k;
        if (mask[id]) {
          mskp[id] = mskp[id - 1] = mskp[id + 1] = 1;
          mskp[id - sx] = mskp[id + sx] = mskp[id - sxy] = mskp[id + sxy] = 1;
        }
      }
  /* main.cu:308-329: mask rank and dilated-mask rank in raster order */
  int cnt = 0, cntp = 0;
  for (int id = 0; id < nsize; id++) {
    host_idx[id] = -1;
    host_idx_sc[id] = -1;
    if (mask[id])
      host_idx[id] = cnt++;
    if (mskp[id])
      host_idx_sc[id] = cntp++;
  }
  /* main.cu:333-394 with cnt_nbrs: per-colour raster lists and in-mask counts */
  int n[2] = {0, 0};
  int *vox[2] = {vox0, vox1};
  unsigned char *nb[2] = {nbrs0, nbrs1};
  int off[6] = {-1, 1, -sx, sx, -sxy, sxy};
  for (int k = 1; k < ndepth + 1; k++)
    for (int j = 1; j < ncol + 1; j++)
      for (int i = 1; i < nrow + 1; i++) {
        int id = i + sx * j + sxy * k;
        if (!mask[id])
          continue;
        int c = colour_of(i, j, k);
        vox[c][n[c]] = id;
        unsigned char m = 0;
        for (int d = 0; d < 6; d++)
          m += mask[id + off[d]] != 0;
        nb[c][n[c]] = m;
        n[c]++;
      }
  if (nbr6) {
    /* sorted order: colour-0 list then colour-1 list */
    int *sorted = (int *)malloc(sizeof(int) * (size_t)nsize);
    for (int id = 0; id < nsize; id++)
      sorted[id] = -1;
    for (int c = 0; c < 2; c++)
      for (int q = 0; q < n[c]; q++)
        sorted[vox[c][q]] = q + (c ? n[0] : 0);
    for (int c = 0; c < 2; c++)
      for (int q = 0; q < n[c]; q++) {
        int s = q + (c ? n[0] : 0);
        for (int d = 0; d < 6; d++) {
          int nid = vox[c][q] + off[d];
          nbr6[(size_t)s * 6 + d] = mask[nid] ? sorted[nid] : -1;
        }
      }
    free(sorted);
  }
  free(mskp);
}

/* ------------------------------------------------------------------------- */
/* Dense FP64 helpers, row-major lower (cholesky.cpp:13-86 semantics).        */
/* ------------------------------------------------------------------------- */
int orc_chol(double *a, int n)
{
  for (int j = 0; j < n; j++) {
    double d = a[j * n + j];
    for (int k = 0; k < j; k++)
      d -= a[j * n + k] * a[j * n + k];
    if (d <= 0.0)
      return 0;
    d = sqrt(d);
    a[j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double s = a[i * n + j];
      for (int k = 0; k < j; k++)
        s -= a[i * n + k] * a[j * n + k];
      a[i * n + j] = s / d;
    }
  }
  return 1;
}

/* h holds L (lower) on entry; on exit the full (L L')^-1 */
void orc_chol_invert(int n, double *h)
{
  double *inv = (double *)calloc((size_t)n * n, sizeof(double));
  double *y = (double *)malloc(sizeof(double) * n);
  for (int col = 0; col < n; col++) {
    for (int i = 0; i < n; i++) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++)
        s -= h[i * n + k] * y[k];
      y[i] = s / h[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double s = y[i];
      for (int k = i + 1; k < n; k++)
        s -= h[k * n + i] * inv[k * n + col];
      inv[i * n + col] = s / h[i * n + i];
    }
  }
  memcpy(h, inv, sizeof(double) * (size_t)n * n);
  free(inv);
  free(y);
}

/* ------------------------------------------------------------------------- */
/* One sweep in the reference layout.                                         */
/* ------------------------------------------------------------------------- */
void orc_update_z(const orc_problem *pb, const float *spat_coef, float b, double eps,
                  uint32_t sweep, int method, double *z_out)
{
  int N = pb->nsubs, P = pb->ncovar, nq = (N + 3) / 4;
  for (int c = 0; c < 2; c++) {
    for (int q = 0; q < pb->n_col[c]; q++) {
      int vox = pb->vox[c][q];
      int id = pb->host_idx[vox], idsc = pb->host_idx_sc[vox];
      uint32_t s = (uint32_t)(q + (c ? pb->n_col[0] : 0));
      double base = (double)b * (double)pb->wm[id];
      for (int g = 0; g < nq; g++) {
        uint32_t r[4];
        orc_draw(pb->key, s * (uint32_t)nq + (uint32_t)g, sweep, ORC_TAG_LATENT, 0, pb->chain, r);
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        if (method == 1) {  /* Box-Muller pairs (u0, u1), (u2, u3) */
          orc_box_muller(orc_u01(r[0]), orc_u01(r[1]), &w[0], &w[1]);
          orc_box_muller(orc_u01(r[2]), orc_u01(r[3]), &w[2], &w[3]);
        }
        for (int l = 0; l < 4; l++) {
          int i = 4 * g + l;
          if (i >= N)
            break;
          double mu = base;
          for (int k = 0; k < P; k++)
            mu += (double)spat_coef[(size_t)k * pb->totvoxp + idsc] * (double)pb->covar[(size_t)k * N + i];
          int y = pb->data[(size_t)i * pb->totvox + id] == 1;
          double z;
          if (method == 1) {
            /* keep the candidate inside the interval of y, else redraw exactly
               from the independent block (FALLBACK, sub = l) */
            double zc = mu + w[l];
            int ok = y ? (zc > eps && zc < 500.0) : (zc < -eps && zc > -500.0);
            if (ok) {
              z = zc;
            } else {
              uint32_t f[4];
              orc_draw(pb->key, s * (uint32_t)nq + (uint32_t)g, sweep, ORC_TAG_FALLBACK, (uint32_t)l, pb->chain, f);
              z = orc_latent_draw(mu, y, orc_u01(f[0]), eps);
            }
          } else {
            z = orc_latent_draw(mu, y, orc_u01(r[l]), eps);
          }
          z_out[(size_t)i * pb->totvox + id] = z;
        }
      }
    }
  }
}

/* covarGPU.cu:327-352: m_j = sum_i x_i (Z_ij - b WM_j), here in FP64 */
void orc_xtz(const orc_problem *pb, const float *z, float b, double *m_out)
{
  int N = pb->nsubs, P = pb->ncovar, V = pb->totvox;
  for (int v = 0; v < V; v++) {
    double bw = (double)b * (double)pb->wm[v];
    for (int k = 0; k < P; k++) {
      double acc = 0.0;
      for (int i = 0; i < N; i++)
        acc += (double)pb->covar[(size_t)k * N + i] * ((double)z[(size_t)i * V + v] - bw);
      m_out[(size_t)k * V + v] = acc;
    }
  }
}

/* covarGPU.cu:588-683 (betaMGPU, betaVGPU): the sufficient statistics of the
 * WM coefficient, over the mask in raster order, in FP64. */
void orc_wm_stat(const orc_problem *pb, const float *spat_coef, const double *z, double *S, double *W)
{
  int N = pb->nsubs, P = pb->ncovar, V = pb->totvox;
  double s = 0.0, w = 0.0;
  for (int c = 0; c < 2; c++) {
    for (int q = 0; q < pb->n_col[c]; q++) {
      int vox = pb->vox[c][q];
      int id = pb->host_idx[vox], idsc = pb->host_idx_sc[vox];
      double wm = (double)pb->wm[id], r = 0.0;
      for (int i = 0; i < N; i++) {
        r += z[(size_t)i * V + id];
        for (int k = 0; k < P; k++)
          r -= (double)spat_coef[(size_t)k * pb->totvoxp + idsc] * (double)pb->covar[(size_t)k * N + i];
      }
      s += wm * r;
      w += wm * wm;
    }
  }
  *S = s;
  *W = (double)N * w;
}

/* mcmc.cpp:1099-1136 (CPU path, NSUBTYPES covariates then the rest: every
 * covariate once): voxels in raster order of the padded grid (k, j, i with i
 * fastest), Kahan-compensated float sums of WM^2 and of WM_j t_j, where t_j is
 * a plain float sum over subjects of Z_ij - sum_c b_c,j x_c,i. */
void orc_cpu_wm_stat(const orc_problem *pb, const float *spat_coef, const float *z, const float *wm,
                     float prior_mean, float prior_prec, float *mean_out, float *var_out)
{
  int N = pb->nsubs, P = pb->ncovar, V = pb->totvox, VP = pb->totvoxp;
  int sx = pb->nrow + 2, sxy = (pb->nrow + 2) * (pb->ncol + 2);
  float mean = 0.0f, var = 0.0f, c = 0.0f, cm = 0.0f;
  for (int k = 1; k <= pb->ndepth; k++)
    for (int j = 1; j <= pb->ncol; j++)
      for (int i = 1; i <= pb->nrow; i++) {
        int idx = i + sx * j + sxy * k;
        int id = pb->host_idx[idx];
        if (id < 0)
          continue;
        int idsc = pb->host_idx_sc[idx];
        float w = wm[id];
        float y = w * w - c;
        float t = var + y;
        c = (t - var) - y;
        var = t;
        float tmp = 0.0f;
        for (int s = 0; s < N; s++) {
          tmp += z[(size_t)s * V + id];
          for (int a = 0; a < P; a++)
            tmp -= spat_coef[(size_t)a * VP + idsc] * pb->covar[(size_t)a * N + s];
        }
        tmp *= w;
        float ym = tmp - cm;
        float tm = mean + ym;
        cm = (tm - mean) - ym;
        mean = tm;
      }
  var *= (float)N;
  var += prior_prec;
  var = 1.0f / var;
  *mean_out = var * (mean + prior_mean * prior_prec);
  *var_out = var;
}

/* covarGPU.cu:711-717: var = 1/W, mean = S var, b = mean + sqrt(var) n */
double orc_wm_draw(double S, double W, const uint32_t key[2], uint32_t sweep, uint32_t chain)
{
  uint32_t r[4];
  double n0, n1;
  orc_draw(key, 0u, sweep, ORC_TAG_BWM, 0, chain, r);
  orc_box_muller(orc_u01(r[0]), orc_u01(r[1]), &n0, &n1);
  double var = 1.0 / W;
  return S * var + sqrt(var) * n0;
}

/* covarGPU.cu:354-392 (probit coefficient update of one colour):
 *   sumV = sum of the 6 dilated-mask neighbours' beta (out-of-mask ones are 0)
 *   mean = m + Prec sumV,  var = X'X + n_j Prec
 *   var <- chol; invert  (Sigma = var^-1);  tmp2 = Sigma mean
 *   beta = tmp2 + chol_lower(Sigma) z       (rmvnormGPU, covarGPU.cu:178-202) */
void orc_update_coef(const orc_problem *pb, int colour, const double *m, const float *prec,
                     double *spat_coef, uint32_t sweep, const double *z_inj)
{
  int P = pb->ncovar, V = pb->totvox, VP = pb->totvoxp;
  int sx = pb->nrow + 2, sxy = (pb->nrow + 2) * (pb->ncol + 2);
  int off[6] = {-1, 1, -sx, sx, -sxy, sxy};
  double *var = (double *)malloc(sizeof(double) * P * P);
  double *mean = (double *)malloc(sizeof(double) * P);
  double *sumv = (double *)malloc(sizeof(double) * P);
  double *tmp2 = (double *)malloc(sizeof(double) * P);
  double *z = (double *)malloc(sizeof(double) * (P + 4));
  for (int q = 0; q < pb->n_col[colour]; q++) {
    int vox = pb->vox[colour][q];
    int id = pb->host_idx[vox], idsc = pb->host_idx_sc[vox];
    uint32_t s = (uint32_t)(q + (colour ? pb->n_col[0] : 0));
    for (int a = 0; a < P; a++) {
      double acc = 0.0;
      for (int d = 0; d < 6; d++)
        acc += spat_coef[(size_t)a * VP + pb->host_idx_sc[vox + off[d]]];
      sumv[a] = acc;
    }
    for (int a = 0; a < P; a++) {
      mean[a] = m[(size_t)a * V + id];
      for (int bb = 0; bb < P; bb++) {
        mean[a] += (double)prec[bb + a * P] * sumv[bb];
        var[a * P + bb] = (double)pb->xxprime[bb + a * P] +
                          (double)pb->nbrs[colour][q] * (double)prec[bb + a * P];
      }
    }
    orc_chol(var, P);
    orc_chol_invert(P, var);
    for (int a = 0; a < P; a++) {
      tmp2[a] = 0.0;
      for (int bb = 0; bb < P; bb++)
        tmp2[a] += var[a * P + bb] * mean[bb];
    }
    orc_chol(var, P);
    if (z_inj) {
      for (int a = 0; a < P; a++)
        z[a] = z_inj[(size_t)q * P + a];
    } else {
      for (int g = 0; 4 * g < P; g++) {
        uint32_t r[4];
        orc_draw(pb->key, s, sweep, ORC_TAG_COEF, (uint32_t)g, pb->chain, r);
        orc_box_muller(orc_u01(r[0]), orc_u01(r[1]), &z[4 * g], &z[4 * g + 1]);
        orc_box_muller(orc_u01(r[2]), orc_u01(r[3]), &z[4 * g + 2], &z[4 * g + 3]);
      }
    }
    for (int a = 0; a < P; a++) {
      double acc = tmp2[a];
      for (int bb = 0; bb <= a; bb++)
        acc += var[a * P + bb] * z[bb];
      spat_coef[(size_t)a * VP + idsc] = acc;
    }
  }
  free(var);
  free(mean);
  free(sumv);
  free(tmp2);
  free(z);
}

/* covarGPU.cu:953-1013 / covar.cpp:220-262: over every mask voxel and its
 * +x, +y, +z neighbour k, add (b^a_j - b^a_k)(b^c_j - b^c_k) when b^a_k != 0.
 * The differences are float (as the reference); products and sums are FP64. */
void orc_pair_sums(const orc_problem *pb, const float *spat_coef, double *t_out)
{
  int P = pb->ncovar, VP = pb->totvoxp;
  int sx = pb->nrow + 2, sxy = (pb->nrow + 2) * (pb->ncol + 2);
  int off[3] = {1, sx, sxy};
  memset(t_out, 0, sizeof(double) * P * P);
  for (int c = 0; c < 2; c++)
    for (int q = 0; q < pb->n_col[c]; q++) {
      int vox = pb->vox[c][q];
      int idsc = pb->host_idx_sc[vox];
      for (int d = 0; d < 3; d++) {
        int nb = pb->host_idx_sc[vox + off[d]];
        for (int a = 0; a < P; a++) {
          float t1 = spat_coef[(size_t)a * VP + nb];
          if (t1 == 0.0f)
            continue;
          float da = spat_coef[(size_t)a * VP + idsc] - t1;
          for (int bb = 0; bb <= a; bb++) {
            float db = spat_coef[(size_t)bb * VP + idsc] - spat_coef[(size_t)bb * VP + nb];
            t_out[a * P + bb] += (double)da * (double)db;
          }
        }
      }
    }
  for (int a = 0; a < P; a++)
    for (int bb = 0; bb < a; bb++)
      t_out[bb * P + a] = t_out[a * P + bb];
}

/* Marsaglia & Tsang (2000), ACM TOMS 26:363-372, for shape a >= 1. Each trial
 * uses one Philox block (row, sweep, GAMMA, trial, chain): a Box-Muller normal
 * from (u0,u1) and the acceptance uniform u2. */
double orc_gamma(double a, const uint32_t key[2], uint32_t row, uint32_t sweep, uint32_t chain)
{
  double d = a - 1.0 / 3.0;
  double c = 1.0 / sqrt(9.0 * d);
  for (uint32_t t = 0; t < 64; t++) {
    uint32_t r[4];
    orc_draw(key, row, sweep, ORC_TAG_GAMMA, t, chain, r);
    double u0 = orc_u01(r[0]), u1 = orc_u01(r[1]), u2 = orc_u01(r[2]);
    double x = sqrt(-2.0 * log(u0)) * cos(2.0 * M_PI * u1);
    double v = 1.0 + c * x;
    if (v <= 0.0)
      continue;
    v = v * v * v;
    if (log(u2) < 0.5 * x * x + d - d * v + d * log(v))
      return d * v;
  }
  return d;
}

/* covarGPU.cu:1086-1108 with rwishart2 (cholesky.cpp:245-286):
 * V = T + I; chol; invert; chol -> L = chol(S), S = (T+I)^-1;
 * Bartlett A (A_ii = sqrt(chi2(df-i)), A_ij ~ N(0,1) for j < i);
 * Lambda = (L A)(L A)'. */
void orc_wishart(int p, const double *t, int df, const uint32_t key[2], uint32_t sweep,
                 uint32_t chain, const double *chi2_inj, const double *norm_inj,
                 double *lambda_out)
{
  double *v = (double *)malloc(sizeof(double) * p * p);
  double *A = (double *)calloc((size_t)p * p, sizeof(double));
  double *B = (double *)calloc((size_t)p * p, sizeof(double));
  for (int i = 0; i < p * p; i++)
    v[i] = t[i];
  for (int i = 0; i < p; i++)
    v[i * p + i] += 1.0;
  orc_chol(v, p);
  orc_chol_invert(p, v);
  orc_chol(v, p);
  for (int i = 0; i < p; i++) {
    double chi2;
    if (chi2_inj)
      chi2 = chi2_inj[i];
    else
      chi2 = 2.0 * orc_gamma(0.5 * (double)(df - i), key, (uint32_t)i, sweep, chain);
    A[i * p + i] = sqrt(chi2);
    for (int j = 0; j < i; j++) {
      if (norm_inj) {
        A[i * p + j] = norm_inj[i * p + j];
      } else {
        uint32_t r[4];
        double z0, z1;
        orc_draw(key, (uint32_t)(i * p + j), sweep, ORC_TAG_WNORM, 0, chain, r);
        orc_box_muller(orc_u01(r[0]), orc_u01(r[1]), &z0, &z1);
        A[i * p + j] = z0;
      }
    }
  }
  for (int i = 0; i < p; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = j; k <= i; k++)
        s += v[i * p + k] * A[k * p + j];
      B[i * p + j] = s;
    }
  for (int i = 0; i < p; i++)
    for (int j = 0; j < p; j++) {
      int K = i < j ? i : j;
      double s = 0.0;
      for (int k = 0; k <= K; k++)
        s += B[i * p + k] * B[j * p + k];
      lambda_out[i * p + j] = s;
    }
  free(v);
  free(A);
  free(B);
}

/* ------------------------------------------------------------------------- */
/* Host KISS (Marsaglia & Zaman) exactly as randgen.cpp:9-23, with the        */
/* reference's scale M = exp(-64 ln 2) (main.cu:54), which is not 2^-64.      */
/* ------------------------------------------------------------------------- */
double orc_kiss(unsigned long *seed)
{
  static double M = 0.0;
  if (M == 0.0)
    M = exp(-64 * log(2.0));
  seed[1] = seed[1] ^ (seed[1] << 17);
  seed[2] = (seed[2] ^ (seed[2] << 18)) & 0x7FFFFFFF;
  seed[0] = 69069 * seed[0] + 23606797;
  seed[1] ^= (seed[1] >> 15);
  seed[2] ^= (seed[2] >> 13);
  return (seed[0] + seed[1] + seed[2]) * M;
}

/* randgen.cpp:36-45 */
unsigned long orc_runiform_long_n(unsigned long n, unsigned long *seed)
{
  unsigned long i = (unsigned long)floorl(n * orc_kiss(seed));
  if (i == n)
    return i - 1;
  return i;
}
