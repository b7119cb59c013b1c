// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Links the reference's CPU-path sources, compiled unchanged from their place
// under the reference tree (bsglmm/covar.cpp, randgen.cpp, cholesky.cpp), into
// oracle/_ref/libbsg_ref.so, and exposes them through a C ABI for the tests
// and for the CPU baseline of bench.py.
//
// The reference program defines the process globals these files read in
// main.cu (NROW..TOTVOXp, INDX, hostIdx, hostIdxSC, M, PREC: main.cu:12-54)
// and read_data.cpp (XXprime: read_data.cpp:321-327). main.cu and mcmc.cpp
// need the CUDA runtime and cannot be built here, so this file defines those
// globals and fills them from arrays the caller passes in, the way main()
// fills them (main.cu:216-394). Nothing of the reference's code is restated
// here; every numeric routine called below is the reference's own.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "binCAR.h"
#include "cholesky.h"
#include "randgen.h"

int PREC = 64;
double M = exp(-PREC * log(2.0));
int NSUBS, NROW, NCOL, NDEPTH, TOTVOX, TOTVOXp, NCOVAR;
int NSUBTYPES = 1;
int MODEL = 1;
int NSIZE;
int iter;
INDEX *INDX;
int *hostIdx, *hostIdxSC;
float *XXprime;
float *XXprime_Fix;

// reference functions (covar.cpp), declared as mcmc.cpp:96-127 declares them
void updateZ(float *, float *, unsigned char *, unsigned char *, float, float *, float *, float *,
             unsigned long *);
void updateSpatCoef(float *, float *, float *, float *, float *, float *, unsigned char *, float,
                    float *, unsigned long *);
void updateSpatCoefPrec(float *, float *, unsigned char *, unsigned long *);
double Spat_Coef_Precision(float *, float *, const int, int *);

static void release()
{
  if (INDX) {
    for (int c = 0; c < 2; c++) {
      free(INDX[c].hostVox);
      free(INDX[c].hostNBRS);
    }
    free(INDX);
    INDX = nullptr;
  }
  free(hostIdx);
  free(hostIdxSC);
  free(XXprime);
  hostIdx = hostIdxSC = nullptr;
  XXprime = nullptr;
}

extern "C" {

int ref_setup(int nrow, int ncol, int ndepth, int nsubs, int ncovar, int totvox, int totvoxp,
              int n0, int n1, const int *idx, const int *idxsc, const int *vox0,
              const int *vox1, const unsigned char *nb0, const unsigned char *nb1,
              const float *xxprime)
{
  release();
  NROW = nrow;
  NCOL = ncol;
  NDEPTH = ndepth;
  NSUBS = nsubs;
  NCOVAR = ncovar;
  TOTVOX = totvox;
  TOTVOXp = totvoxp;
  NSIZE = (nrow + 2) * (ncol + 2) * (ndepth + 2);
  hostIdx = (int *)malloc(sizeof(int) * NSIZE);
  hostIdxSC = (int *)malloc(sizeof(int) * NSIZE);
  memcpy(hostIdx, idx, sizeof(int) * NSIZE);
  memcpy(hostIdxSC, idxsc, sizeof(int) * NSIZE);
  INDX = (INDEX *)calloc(2, sizeof(INDEX));
  const int *vx[2] = {vox0, vox1};
  const unsigned char *nb[2] = {nb0, nb1};
  int n[2] = {n0, n1};
  for (int c = 0; c < 2; c++) {
    INDX[c].hostN = n[c];
    INDX[c].hostVox = (int *)malloc(sizeof(int) * (n[c] ? n[c] : 1));
    INDX[c].hostNBRS = (unsigned char *)malloc(n[c] ? n[c] : 1);
    memcpy(INDX[c].hostVox, vx[c], sizeof(int) * n[c]);
    memcpy(INDX[c].hostNBRS, nb[c], n[c]);
  }
  XXprime = (float *)malloc(sizeof(float) * ncovar * ncovar);
  memcpy(XXprime, xxprime, sizeof(float) * ncovar * ncovar);
  return 0;
}

double ref_kiss(unsigned long *seed) { return kiss(seed); }
double ref_snorm(unsigned long *seed) { return snorm(seed); }
double ref_rgamma(double a, double b, unsigned long *seed) { return rgamma(a, b, seed); }
double ref_truncnorm2(double mu, double sd, double u, double v, unsigned long *seed)
{
  return truncNorm2(mu, sd, u, v, seed);
}
unsigned long ref_runiform_long_n(unsigned long n, unsigned long *seed)
{
  return runiform_long_n(n, seed);
}
int ref_rwishart2(double *result, double *S, int p, int df, unsigned long *seed)
{
  return rwishart2(result, S, p, df, seed);
}
int ref_cholesky_decomp(double *a, int n) { return cholesky_decomp(a, n); }
void ref_cholesky_invert(int n, double *h) { cholesky_invert(n, h); }

// updateZ(Z, alpha, data, msk, beta, WM, spatCoef, covar, seed): alpha and msk unused
void ref_update_z(float *Z, unsigned char *data, float beta, float *WM, float *spatCoef,
                  float *covar, unsigned long *seed)
{
  updateZ(Z, nullptr, data, nullptr, beta, WM, spatCoef, covar, seed);
}

void ref_update_spat_coef(float *covar, float *spatCoef, float *prec, float *Z, float beta,
                          float *WM, unsigned long *seed)
{
  float alpha_mean[32];
  updateSpatCoef(covar, spatCoef, prec, nullptr, alpha_mean, Z, nullptr, beta, WM, seed);
}

void ref_update_spat_coef_prec(float *prec, float *spatCoef, unsigned long *seed)
{
  updateSpatCoefPrec(prec, spatCoef, nullptr, seed);
}

double ref_spat_coef_precision(float *sc1, float *sc2, int colour)
{
  return Spat_Coef_Precision(sc1, sc2, INDX[colour].hostN, INDX[colour].hostVox);
}

// One CPU-path sweep in the order of mcmc.cpp:400-428, without the WM
// coefficient update (mcmc.cpp:409 lives in mcmc.cpp, which needs CUDA).
void ref_sweep(float *Z, unsigned char *data, float *WM, float *spatCoef, float *covar,
               float *prec, unsigned long *seed)
{
  ref_update_z(Z, data, 0.0f, WM, spatCoef, covar, seed);
  ref_update_spat_coef(covar, spatCoef, prec, Z, 0.0f, WM, seed);
  ref_update_spat_coef_prec(prec, spatCoef, seed);
}

}  // extern "C"
