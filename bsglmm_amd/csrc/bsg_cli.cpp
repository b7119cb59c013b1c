// bsg_cli.cpp -- the BinCar command line on top of libbsglmm (bsg_bincar_main).
//
// Same positional arguments, ./images resolution, input type rules, seed.dat
// handling, console report and output files as the reference program
// (main.cu:60-166, 409-432; mcmc.cpp:300-975; read_data.cpp:187-545). The
// sampler runs as one chain through the C ABI; the per-save work (posterior
// sums, lesion probabilities, the per-subject log-likelihood behind DIC and
// Qhat) stays on the device and only N * NTypes doubles come back per save.
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/bsglmm.h"
#include "../../include/bsglmm_io.h"

namespace {

const char *IMAGE_DIR = "./images";
const char *DEFAULT_MASK = "mask.nii.gz";
const char *DEFAULT_WM = "avg152T1_white.nii.gz";

void usage(const char *prog)
{
  printf("Usage: %s  NTypes  NCov  GPU  Design Mask WM [MaxIter BurnIn]  \n", prog);
  printf("  NTypes  - Number of groups\n");
  printf("  NCov    - Number of covariates (count must include groups)\n");
  printf("  GPU     - 1 use GPU (this build always samples on the GPU)\n");
  printf("  Design  - Text file, tab or space separated data file\n");
  printf("  Mask    - Filename of mask image in '%s' directory, or '1' to use default (%s)\n", IMAGE_DIR,
         DEFAULT_MASK);
  printf("  WM      - Filename of WM image in '%s' directory, '0' to use none, or '1' to use default (%s)\n",
         IMAGE_DIR, DEFAULT_WM);
  printf("  MaxIter - Number of iterations (defaults to 100,000)\n");
  printf("  BurnIn  - Number of burn-in iterations (defaults to 50,000)\n");
  printf("Options after these (this build):\n");
  printf("  -checkpoint K  every K iterations write last_state.bsg and last_run.dat (exact state) and\n");
  printf("                 last_parms.dat / last_spatCoef.dat (the reference's save_iter files)\n");
  printf("  -stop S        stop after iteration S, writing the checkpoint\n");
  printf("  -restart       continue the run saved in last_state.bsg / last_run.dat\n");
}

// ---- checkpoints of a run (save_iter / restart_iter, mcmc.cpp:1265-1336) ----
// last_state.bsg holds the library's chain state (bsg_checkpoint_write);
// last_run.dat this program's own counters: the next iteration, the number of
// saves, the DIC and Qhat sums, the run's iteration counts and the seed.dat
// triple it will write at the end.
const char *CKPT_STATE = "last_state.bsg";
const char *CKPT_RUN = "last_run.dat";
const char RUN_MAGIC[8] = {'B', 'S', 'G', 'R', 'U', 'N', '0', '1'};

struct RunState {
  int next_iter = 0, totsav = 0, maxiter = 0, burnin = 0, nsubs = 0, ntypes = 0;
  double ED = 0.0;
  unsigned long seed_out[3] = {0, 0, 0};
  std::vector<std::vector<double>> Qhat;
};

bool write_run_state(const RunState &st)
{
  FILE *f = fopen(CKPT_RUN, "wb");
  if (!f) return false;
  int32_t hdr[6] = {st.next_iter, st.totsav, st.maxiter, st.burnin, st.nsubs, st.ntypes};
  bool ok = fwrite(RUN_MAGIC, 1, 8, f) == 8 && fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
            fwrite(&st.ED, sizeof(double), 1, f) == 1 && fwrite(st.seed_out, sizeof(st.seed_out), 1, f) == 1;
  for (int i = 0; ok && i < st.nsubs; i++)
    ok = fwrite(st.Qhat[i].data(), sizeof(double), st.ntypes, f) == (size_t)st.ntypes;
  return (fclose(f) == 0) && ok;
}

bool read_run_state(RunState &st)
{
  FILE *f = fopen(CKPT_RUN, "rb");
  if (!f) return false;
  char magic[8];
  int32_t hdr[6];
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, RUN_MAGIC, 8) == 0 && fread(hdr, sizeof(hdr), 1, f) == 1 &&
            fread(&st.ED, sizeof(double), 1, f) == 1 && fread(st.seed_out, sizeof(st.seed_out), 1, f) == 1;
  if (ok) {
    st.next_iter = hdr[0];
    st.totsav = hdr[1];
    st.maxiter = hdr[2];
    st.burnin = hdr[3];
    st.nsubs = hdr[4];
    st.ntypes = hdr[5];
    ok = st.nsubs > 0 && st.ntypes > 0 && st.nsubs <= (1 << 24) && st.ntypes <= BSG_MAX_COVAR;
  }
  if (ok) st.Qhat.assign(st.nsubs, std::vector<double>(st.ntypes, 0.0));
  for (int i = 0; ok && i < st.nsubs; i++)
    ok = fread(st.Qhat[i].data(), sizeof(double), st.ntypes, f) == (size_t)st.ntypes;
  fclose(f);
  return ok;
}

// Failures of the library print its message and end the program with the
// reference's exit status for that kind of failure.
#define LIBCHK(x, code)                                              \
  do {                                                               \
    if ((x) != BSG_OK) {                                             \
      printf("%s\n", bsg_last_error());                              \
      return (code);                                                 \
    }                                                                \
  } while (0)

// Padded raster index (main.cu:56)
inline size_t pidx(int i, int j, int k, int nr, int nc) { return i + (size_t)(nr + 2) * (j + (size_t)(nc + 2) * k); }

// Cholesky inverse of a small SPD matrix in place (cholesky.cpp:13-86 use in mcmc.cpp:551-552)
bool spd_inverse(std::vector<double> &a, int n)
{
  std::vector<double> L((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    double d = a[j * n + j];
    for (int k = 0; k < j; k++) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0)) return false;
    L[j * n + j] = sqrt(d);
    for (int i = j + 1; i < n; i++) {
      double s = a[i * n + j];
      for (int k = 0; k < j; k++) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / L[j * n + j];
    }
  }
  std::vector<double> inv((size_t)n * n, 0.0), y(n);
  for (int c = 0; c < n; c++) {
    for (int i = 0; i < n; i++) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) s -= L[i * n + k] * y[k];
      y[i] = s / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double s = y[i];
      for (int k = i + 1; k < n; k++) s -= L[k * n + i] * inv[k * n + c];
      inv[i * n + c] = s / L[i * n + i];
    }
  }
  a = inv;
  return true;
}

// ProbSDNorm (covar.cpp:327-348), A&S 26.2.17 in double
double prob_sdnorm(double x)
{
  const double t = 1.0 / (1.0 + 0.2316419 * fabs(x));
  const double poly = t * (0.319381530 + t * (-0.356563782 + t * (1.781477937 + t * (-1.821255978 + t * 1.330274429))));
  const double tail = 0.39894228 * exp(-x * x / 2.0) * poly;
  return x >= 0.0 ? 1.0 - tail : tail;
}

struct Run {
  int ntypes = 0, ncovar = 0, nsubs = 0;
  int nrow = 0, ncol = 0, ndepth = 0, totvox = 0;
  bool update_wm = true;
  std::vector<unsigned char> msk;  // padded mask
  std::vector<int> rank_grid;      // unpadded raster index -> mask rank, -1 outside
  std::vector<float> covar, wm;
  std::vector<unsigned char> data;
  std::vector<std::string> names;
  std::vector<double> grp_prior;
};

// read_nifti1_mask (read_data.cpp:187-247): FLOAT64 required
int read_mask(const std::string &name, Run &r)
{
  int dims[3], dt;
  if (bsg_nifti_info(name.c_str(), dims, &dt) != BSG_OK) {
    printf("%s\n", bsg_last_error());
    return 1;
  }
  if (dt != BSG_NIFTI_FLOAT64) {
    printf("Incorrect data type in file: %d\n", dt);
    printf("Must be float (NIFTI_TYPE_FLOAT64)\n");
    return 0;
  }
  r.nrow = dims[0];
  r.ncol = dims[1];
  r.ndepth = dims[2];
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  std::vector<double> img(n);
  if (bsg_nifti_read(name.c_str(), img.data(), n) != BSG_OK) {
    printf("\n%s\n", bsg_last_error());
    return 0;
  }
  r.msk.assign((size_t)(r.nrow + 2) * (r.ncol + 2) * (r.ndepth + 2), 0);
  r.rank_grid.assign(n, -1);
  size_t t = 0;
  r.totvox = 0;
  for (int k = 1; k <= r.ndepth; k++)
    for (int j = 1; j <= r.ncol; j++)
      for (int i = 1; i <= r.nrow; i++, t++) {
        const unsigned char v = (unsigned char)img[t];
        r.msk[pidx(i, j, k, r.nrow, r.ncol)] = v;
        if (v) r.rank_grid[t] = r.totvox++;
      }
  return -1;  // success
}

// read_nifti1_WM (read_data.cpp:489-545): UINT8 required, WM = value / 255
int read_wm(const std::string &name, Run &r)
{
  r.wm.assign(r.totvox, 1.0f);
  if (!r.update_wm) return -1;
  int dims[3], dt;
  if (bsg_nifti_info(name.c_str(), dims, &dt) != BSG_OK) {
    printf("%s\n", bsg_last_error());
    return 1;
  }
  if (dt != BSG_NIFTI_UINT8) {
    printf("Incorrect data type in file: %d\n", dt);
    printf("Must be unsigned char (NIFTI_TYPE_UINT8)\n");
    return 0;
  }
  const size_t n = (size_t)r.nrow * r.ncol * r.ndepth;
  if ((size_t)dims[0] * dims[1] * dims[2] != n) {
    printf("\nError reading volume 1 from %s: size differs from the mask\n", name.c_str());
    return 0;
  }
  std::vector<double> img(n);
  if (bsg_nifti_read(name.c_str(), img.data(), n) != BSG_OK) {
    printf("\n%s\n", bsg_last_error());
    return 0;
  }
  for (size_t t = 0; t < n; t++)
    if (r.rank_grid[t] >= 0) r.wm[r.rank_grid[t]] = (float)(unsigned char)img[t] / 255.0f;
  return -1;
}

// read_nifti1_image (read_data.cpp:249-416): design file, then one FLOAT32 or
// FLOAT64 lesion image per subject, cast to unsigned char at mask voxels
int read_images(const char *design, Run &r)
{
  int nsubs = 0;
  if (bsg_design_read(design, r.ncovar, &nsubs, nullptr, nullptr, nullptr) != BSG_OK) {
    printf("%s\n", bsg_last_error());
    return 1;
  }
  r.nsubs = nsubs;
  r.covar.assign((size_t)r.ncovar * nsubs, 0.0f);
  std::vector<char> names((size_t)r.ncovar * 64), images((size_t)nsubs * 512);
  if (bsg_design_read(design, r.ncovar, &nsubs, r.covar.data(), names.data(), images.data()) != BSG_OK) {
    printf("%s\n", bsg_last_error());
    return 1;
  }
  for (int c = 0; c < r.ncovar; c++) r.names.emplace_back(&names[64 * (size_t)c]);
  r.grp_prior.assign(r.ntypes, 0.0);
  for (int i = 0; i < nsubs; i++)
    for (int t = 0; t < r.ntypes; t++) r.grp_prior[t] += (int)r.covar[(size_t)t * nsubs + i];
  for (int t = 0; t < r.ntypes; t++) r.grp_prior[t] /= (double)nsubs;
  const size_t n = (size_t)r.nrow * r.ncol * r.ndepth;
  r.data.assign((size_t)nsubs * r.totvox, 0);
  std::vector<double> img(n);
  fflush(stdout);
  fprintf(stderr, "Loading images: ");
  for (int i = 0; i < nsubs; i++) {
    const char *nm = &images[512 * (size_t)i];
    fprintf(stderr, ".");
    int dims[3], dt;
    if (bsg_nifti_info(nm, dims, &dt) != BSG_OK) {
      printf("%s\n", bsg_last_error());
      return 1;
    }
    if (dt != BSG_NIFTI_FLOAT32 && dt != BSG_NIFTI_FLOAT64) {
      printf("Incorrect data type in file %s: %d\n", nm, dt);
      printf("Must be float (NIFTI_TYPE_FLOAT32) or double (NIFTI_TYPE_FLOAT64)\n");
      return 0;
    }
    if ((size_t)dims[0] * dims[1] * dims[2] != n || bsg_nifti_read(nm, img.data(), n) != BSG_OK) {
      printf("\nError reading volume 1 from %s\n", nm);
      return 0;
    }
    for (size_t t = 0; t < n; t++)
      if (r.rank_grid[t] >= 0) {
        // FLOAT32 images are cast through float, as read_data.cpp:380
        const double v = dt == BSG_NIFTI_FLOAT32 ? (double)(float)img[t] : img[t];
        r.data[(size_t)i * r.totvox + r.rank_grid[t]] = (unsigned char)v;
      }
  }
  fprintf(stderr, "\n");
  return -1;
}

// Scatters a mask-rank vector to the unpadded grid (x fastest), 0 outside
std::vector<double> to_grid(const Run &r, const double *v)
{
  std::vector<double> g(r.rank_grid.size(), 0.0);
  for (size_t t = 0; t < g.size(); t++)
    if (r.rank_grid[t] >= 0) g[t] = v[r.rank_grid[t]];
  return g;
}

int write_map(const Run &r, const std::string &name, const std::vector<double> &g)
{
  if (bsg_nifti_write_f64((name + ".gz").c_str(), r.nrow, r.ncol, r.ndepth, g.data()) != BSG_OK) {
    fprintf(stderr, "\n%s\n", bsg_last_error());
    return 1;
  }
  return 0;
}

// write_empir_prb (read_data.cpp:418-487)
int write_empir_prb(const Run &r)
{
  const size_t n = r.rank_grid.size();
  std::vector<double> sum(n, 0.0), sum2(n, 0.0);
  for (int t = 0; t < r.ntypes; t++) {
    double N = 0.0;
    for (int i = 0; i < r.nsubs; i++)
      if (r.covar[(size_t)t * r.nsubs + i] != 0.0f) N++;
    for (size_t g = 0; g < n; g++) {
      sum[g] = 0.0;
      const int v = r.rank_grid[g];
      if (v < 0) continue;
      for (int i = 0; i < r.nsubs; i++)
        if (r.covar[(size_t)t * r.nsubs + i] != 0.0f) sum[g] += r.data[(size_t)i * r.totvox + v];
      sum2[g] += sum[g];
      sum[g] /= N;
    }
    if (write_map(r, "empir_prb_" + r.names[t] + ".nii", sum)) return 1;
  }
  if (write_map(r, "total_lesion_cnt.nii", sum2)) return 1;
  for (auto &x : sum2) x /= r.nsubs;
  return write_map(r, "total_empir_prb.nii", sum2);
}

int subtype_of(const Run &r, int i)
{
  for (int t = 0; t < r.ntypes; t++)
    if (r.covar[(size_t)t * r.nsubs + i] != 0.0f) return t;
  return 0;
}

// Qhat.dat / Qhat2.dat rows (mcmc.cpp:719-786)
void write_qhat(const Run &r, const char *path, const std::vector<std::vector<double>> &Q)
{
  FILE *out = fopen(path, "w");
  if (!out) {
    fprintf(stderr, "cannot write %s\n", path);
    return;
  }
  for (int i = 0; i < r.nsubs; i++) {
    fprintf(out, "%d ", i);
    double mx = -DBL_MAX;
    for (int t = 0; t < r.ntypes; t++) mx = (mx < Q[i][t]) ? Q[i][t] : mx;
    double qh = exp(Q[i][0] - mx);
    for (int t = 1; t < r.ntypes; t++) qh += exp(Q[i][t] - mx);
    qh = log(qh) + mx;
    for (int t = 0; t < r.ntypes; t++) fprintf(out, "%.6lf ", exp(Q[i][t] - qh));
    mx = -DBL_MAX;
    int pred = -1;
    for (int t = 0; t < r.ntypes; t++)
      if (mx <= Q[i][t]) {
        mx = Q[i][t];
        pred = t;
      }
    fprintf(out, "%d ", pred);
    fprintf(out, "%d\n", subtype_of(r, i));
  }
  fclose(out);
}

}  // namespace

extern "C" int bsg_bincar_main(int argc, char **argv)
{
  // positional arguments (main.cu:82-112), then this build's options
  int npos = 1;
  while (npos < argc && !(argv[npos][0] == '-' && argv[npos][1] >= 'a' && argv[npos][1] <= 'z')) npos++;
  int ckpt_every = 0, stop_at = -1;
  bool restart = false;
  for (int k = npos; k < argc; k++) {
    if (strcmp(argv[k], "-checkpoint") == 0 && k + 1 < argc) {
      ckpt_every = atoi(argv[++k]);
    } else if (strcmp(argv[k], "-stop") == 0 && k + 1 < argc) {
      stop_at = atoi(argv[++k]);
    } else if (strcmp(argv[k], "-restart") == 0) {
      restart = true;
    } else {
      usage(argv[0]);
      return 1;
    }
  }
  if (ckpt_every < 0) {
    usage(argv[0]);
    return 1;
  }
  argc = npos;
  if (argc != 7 && argc != 9) {
    usage(argv[0]);
    return 1;
  }
  Run r;
  r.ntypes = atoi(argv[1]);
  r.ncovar = atoi(argv[2]);
  const int gpu = atoi(argv[3]);
  std::string mask_name = std::string(IMAGE_DIR) + "/" + (strcmp(argv[5], "1") == 0 ? DEFAULT_MASK : argv[5]);
  std::string wm_name;
  if (strcmp(argv[6], "0") == 0)
    r.update_wm = false;
  else
    wm_name = std::string(IMAGE_DIR) + "/" + (strcmp(argv[6], "1") == 0 ? DEFAULT_WM : argv[6]);
  int maxiter = 100000, burnin = 50000;
  if (argc > 7) {
    maxiter = atoi(argv[7]);
    burnin = atoi(argv[8]);
    if (maxiter < 1000) printf("WARNING: Silly-small number of iterations used; recommend abort and use more\n");
    if (burnin > maxiter) printf("WARNING: Burn-in exceeds MAXITER, no results will be saved\n");
  }
  if (r.ntypes < 1 || r.ncovar < r.ntypes || r.ncovar > BSG_MAX_COVAR) {
    printf("NTypes must be >= 1 and NCov in NTypes..%d\n", BSG_MAX_COVAR);
    return 1;
  }
  if (!gpu) printf("GPU = 0 asks for the reference's CPU sampler; this build samples on the GPU\n");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    printf("There are no available device(s) that support HIP\n");
    return 1;
  }
  printf("Detected %d HIP device(s)\n", ndev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess) printf("Device used is %s (%s)\n\n", prop.name, prop.gcnArchName);

  unsigned long seed[3];
  FILE *fseed = fopen("seed.dat", "r");
  if (!fseed) {
    printf("No seed.dat found; using [ 1 2 3 ]\n");  // message and values as main.cu:157-158
    seed[0] = 1UL;
    seed[1] = 2UL;
    seed[2] = 2UL;
  } else {
    const int rtn = fscanf(fseed, "%lu %lu %lu\n", &seed[0], &seed[1], &seed[2]);
    if (rtn == 0 || fclose(fseed) != 0) return 1;
  }

  int rc;
  if ((rc = read_mask(mask_name, r)) >= 0) return rc;
  if ((rc = read_wm(wm_name, r)) >= 0) return rc;
  if ((rc = read_images(argv[4], r)) >= 0) return rc;
  const int N = r.nsubs, P = r.ncovar, T = r.ntypes, V = r.totvox;

  bsg_layout *lay = nullptr;
  LIBCHK(bsg_layout_create(r.msk.data(), r.nrow, r.ncol, r.ndepth, &lay), 1);
  int lcounts[4];
  LIBCHK(bsg_layout_counts(lay, lcounts), 1);
  const int totvoxp = lcounts[1];
  bsg_opts o;
  bsg_opts_default(&o);
  o.seed = bsg_devseed(seed);  // main.cu:413
  o.xtz_mode = 2;
  o.z_store = 2;
  if (write_empir_prb(r)) return 1;
  bsg_ctx *ctx = nullptr;
  LIBCHK(bsg_create(lay, N, P, r.covar.data(), r.data.data(), r.update_wm ? r.wm.data() : nullptr, &o, &ctx), 1);
  bsg_layout_destroy(lay);

  // mcmc.cpp:303-309
  const int save_iter = (maxiter - burnin) < 10000 ? maxiter - burnin : (maxiter - burnin) / 10000;
  const int prn_iter = 100;
  int totsav_expected = 0;
  if (save_iter > 0)
    for (int it = burnin + 1; it <= maxiter; it++)
      if (it % save_iter == 0) totsav_expected++;
  LIBCHK(bsg_accum_reset(ctx, T, 50, totsav_expected >= 50 ? totsav_expected / 50 : 1), 1);

  std::vector<std::vector<double>> Qhat(N, std::vector<double>(T, 0.0));
  std::vector<double> ll((size_t)N * T), means(P), prec((size_t)P * P);
  std::vector<float> precf((size_t)P * P);
  double ED = 0.0;
  int totsav = 0;
  int iter = 0;
  if (restart) {  // restart_iter (mcmc.cpp:1298-1336), exact
    RunState st;
    if (!read_run_state(st) || st.nsubs != N || st.ntypes != T) {
      printf("-restart: %s is missing or is not a checkpoint of this study\n", CKPT_RUN);
      return 1;
    }
    if (st.maxiter != maxiter || st.burnin != burnin) {
      printf("-restart: the checkpoint is of a run with MaxIter %d and BurnIn %d\n", st.maxiter, st.burnin);
      return 1;
    }
    LIBCHK(bsg_checkpoint_read(ctx, CKPT_STATE), 1);
    iter = st.next_iter;
    totsav = st.totsav;
    ED = st.ED;
    Qhat = st.Qhat;
    for (int k = 0; k < 3; k++) seed[k] = st.seed_out[k];
    printf("RESTART at iteration %d (%d saves so far)\n", iter, totsav);
  }
  // writes the checkpoint files of the run, the next iteration being `next_iter`
  auto checkpoint = [&](int next_iter) -> int {
    LIBCHK(bsg_checkpoint_write(ctx, CKPT_STATE), 1);
    RunState st;
    st.next_iter = next_iter;
    st.totsav = totsav;
    st.maxiter = maxiter;
    st.burnin = burnin;
    st.nsubs = N;
    st.ntypes = T;
    st.ED = ED;
    for (int k = 0; k < 3; k++) st.seed_out[k] = seed[k];
    st.Qhat = Qhat;
    if (!write_run_state(st)) {
      printf("cannot write %s\n", CKPT_RUN);
      return 1;
    }
    // the reference's own files (save_iter, mcmc.cpp:1265-1286): b, the plane
    // means, Lambda as text; the coefficients as float[NCOV][TOTVOXp]
    float b = 0.0f;
    LIBCHK(bsg_get_wm_coef(ctx, 0, &b), 1);
    LIBCHK(bsg_coef_means(ctx, 0, means.data()), 1);
    LIBCHK(bsg_get_prec(ctx, 0, precf.data()), 1);
    FILE *fp = fopen("last_parms.dat", "w");
    if (fp) {
      fprintf(fp, "%f ", b);
      for (int c = 0; c < P; c++) fprintf(fp, "%f ", means[c]);
      for (int a = 0; a < P; a++)
        for (int c = 0; c < P; c++) fprintf(fp, "%f ", precf[c + a * P]);
      fprintf(fp, "\n");
      fclose(fp);
    }
    std::vector<float> sc((size_t)P * totvoxp);
    LIBCHK(bsg_get_spat_coef(ctx, 0, sc.data()), 1);
    fp = fopen("last_spatCoef.dat", "wb");
    if (fp) {
      fwrite(sc.data(), sizeof(float), sc.size(), fp);
      fclose(fp);
    }
    return 0;
  };
  const auto t_start = std::chrono::steady_clock::now();
  const int iter0 = iter;
  auto is_ckpt = [&](int it) { return (ckpt_every > 0 && it > 0 && it % ckpt_every == 0) || it == stop_at; };
  while (iter <= maxiter) {
    // iterations up to the next report, save or checkpoint run as one batch of sweeps
    int next = iter;
    while (next <= maxiter && next % prn_iter != 0 && !(next > burnin && save_iter > 0 && next % save_iter == 0) &&
           !is_ckpt(next))
      next++;
    if (next > iter) {
      LIBCHK(bsg_sweep(ctx, next - iter), 1);
      iter = next;
      continue;
    }
    const bool report = iter % prn_iter == 0;
    if (report) {
      double ms[6];
      LIBCHK(bsg_sweep_profiled(ctx, 1, ms), 1);
      printf("Time to updateZGPU kernel: %f ms\n", ms[0]);
      if (r.update_wm) printf("Time to updateBetaGPU kernel: %f ms\n", ms[1]);
      printf("Time to updateSpatCoefGPU kernel: %f ms\n", ms[2] + ms[3] + (r.update_wm ? 0.0 : ms[1]));
      printf("Time to updateSpatCoefPrecGPU kernel: %f ms\n\n", ms[4] + ms[5]);
    } else {
      LIBCHK(bsg_sweep(ctx, 1), 1);
    }
    if (report) {  // mcmc.cpp:535-582
      float b = 0.0f;
      LIBCHK(bsg_get_wm_coef(ctx, 0, &b), 1);
      LIBCHK(bsg_coef_means(ctx, 0, means.data()), 1);
      LIBCHK(bsg_get_prec(ctx, 0, precf.data()), 1);
      for (int k = 0; k < P * P; k++) prec[k] = precf[k];
      const bool ok = spd_inverse(prec, P);
      printf("iter = %d\t\n", iter);
      printf("  WM mask coefficient \n");
      printf("\t %f \n\n", b);
      printf("  overall mean and SD for groups \n");
      for (int c = 0; c < T; c++) printf("\t %10.6lf %10.6f\n", means[c], ok ? sqrt(prec[c + c * P]) : NAN);
      printf("\n");
      printf("  overall mean and SD for spatially varying coefficients for all covariates \n");
      for (int c = T; c < P; c++) printf("\t %10.6lf %10.6f\n", means[c], ok ? sqrt(prec[c + c * P]) : NAN);
      printf("\n");
      printf("  correlation matrix of regression parameters \n");
      for (int a = 0; a < P; a++) {
        for (int c = 0; c <= a; c++)
          printf("   %6.3lf ", ok ? prec[c + a * P] / sqrt(prec[a + a * P] * prec[c + c * P]) : NAN);
        printf("\n");
      }
      printf("\n");
      fflush(nullptr);
    }
    if (iter > burnin && save_iter > 0 && iter % save_iter == 0) {  // mcmc.cpp:584-651
      const auto t0 = std::chrono::steady_clock::now();
      LIBCHK(bsg_loglik(ctx, 0, T, ll.data()), 1);
      for (int i = 0; i < N; i++) {
        const int st = subtype_of(r, i);
        const double* l = &ll[(size_t)i * T];
        ED += l[st];
        // covarGPU.cu:1935-1952, including its first-save quirk for t >= 1
        for (int t = (totsav == 0 ? 1 : 0); t < T; t++) {
          const double d = l[t] - l[st];
          const double mx = d > Qhat[i][t] ? d : Qhat[i][t];
          Qhat[i][t] = log(exp(Qhat[i][t] - mx) + exp(d - mx)) + mx;
        }
        if (totsav == 0) Qhat[i][0] = l[0] - l[st];
      }
      LIBCHK(bsg_accum_add(ctx), 1);
      totsav++;
      if (report)
        printf("Time to compute_predictGPU: %f ms\n\n",
               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (is_ckpt(iter)) {
      LIBCHK(bsg_synchronize(ctx), 1);
      if (checkpoint(iter + 1)) return 1;
      if (iter == stop_at) {
        printf("stopped after iteration %d; checkpoint in %s and %s (continue with -restart)\n", iter, CKPT_STATE,
               CKPT_RUN);
        bsg_destroy(ctx);
        return 0;
      }
    }
    iter++;
  }
  LIBCHK(bsg_synchronize(ctx), 1);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  printf("TOTSAV = %d\n", totsav);
  printf("%d sweeps in %.2f s (%.1f sweeps/s)\n", maxiter + 1 - iter0, secs, (maxiter + 1 - iter0) / secs);

  write_qhat(r, "Qhat.dat", Qhat);
  {
    auto Q2 = Qhat;
    for (int i = 0; i < N; i++)
      for (int t = 0; t < T; t++) Q2[i][t] = r.grp_prior[t] * Q2[i][t];
    write_qhat(r, "Qhat2.dat", Q2);
  }

  // posterior summaries (mcmc.cpp:795-964)
  std::vector<double> mean((size_t)P * V, NAN), var((size_t)P * V, NAN), prb((size_t)T * V, NAN), bwm(V, NAN);
  int ndraws = 0;
  if (totsav > 0) LIBCHK(bsg_accum_get(ctx, 0, &ndraws, mean.data(), var.data(), prb.data(), bwm.data(), nullptr), 1);
  if (r.update_wm && write_map(r, "bWM.nii", to_grid(r, bwm.data()))) return 1;
  for (int c = 0; c < P; c++) {
    const double *m = &mean[(size_t)c * V], *v = &var[(size_t)c * V];
    std::vector<double> sc(V, 0.0);
    for (int j = 0; j < V; j++)
      if (v[j] > 0.0) sc[j] = m[j] / sqrt(v[j]);
    if (write_map(r, "spatCoef_" + r.names[c] + ".nii", to_grid(r, m)) ||
        write_map(r, "spatCoef_" + r.names[c] + ".Var.nii", to_grid(r, v)) ||
        write_map(r, "standCoef_" + r.names[c] + ".nii", to_grid(r, sc.data())))
      return 1;
  }
  for (int t = 0; t < T; t++)
    if (write_map(r, "prb_" + r.names[t] + ".nii", to_grid(r, &prb[(size_t)t * V]))) return 1;

  // DIC (mcmc.cpp:966-975; compute_prb_DIC, covar.cpp:441-486)
  double DE = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < V; j++) {
      double y = 0.0;
      for (int c = 0; c < P; c++) y += mean[(size_t)c * V + j] * (double)r.covar[(size_t)c * N + i];
      const double x = prob_sdnorm((double)(float)(y + (r.update_wm ? bwm[j] : 0.0)));
      const int d = r.data[(size_t)i * V + j];
      DE += d * log(x + DBL_MIN) + (1 - d) * log(1.0 - x + DBL_MIN);
    }
  DE *= -2.0;
  ED = -2.0 * ED / (double)totsav;
  const double PD = ED - DE;
  printf("DE = %lf ED = %lf PD = %lf DIC = %lf\n", DE, ED, PD, DE + 2 * PD);
  FILE *fdic = fopen("DIC.dat", "w");
  if (fdic) {
    fprintf(fdic, "DE = %lf ED = %lf PD = %lf DIC = %lf\n", DE, ED, PD, DE + 2 * PD);
    fclose(fdic);
  }
  bsg_destroy(ctx);

  fseed = fopen("seed.dat", "w");  // main.cu:430-432
  if (fseed) {
    fprintf(fseed, "%lu %lu %lu\n", seed[0], seed[1], seed[2]);
    fclose(fseed);
  }
  return 0;
}
