/*
 * bsglmm_io.h -- file formats of the reference program, in libbsglmm.so:
 * NIfTI-1 single-file images (.nii / .nii.gz) and the design file.
 *
 * Each entry point names the reference code whose behaviour it keeps. Status
 * codes and bsg_last_error() as in bsglmm.h; nothing exits.
 */
#ifndef BSGLMM_IO_H
#define BSGLMM_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* NIfTI-1 datatype codes (nifti1.h) */
#define BSG_NIFTI_UINT8 2
#define BSG_NIFTI_INT16 4
#define BSG_NIFTI_INT32 8
#define BSG_NIFTI_FLOAT32 16
#define BSG_NIFTI_FLOAT64 64

/* File-name rule of nifti_expand (read_data.cpp:81-140): the name must end in
 * .nii or .nii.gz; when that file is missing, the other spelling is tried.
 * resolved (cap bytes, may be NULL) receives the path that exists. Compressed
 * files are read in process (zlib) instead of the reference's in-place
 * gunzip / gzip round trip, so input files are never rewritten. */
int bsg_nifti_resolve(const char *name, char *resolved, size_t cap);

/* Header of volume 1: dims = dim[1..3], datatype code (nifti1.h). */
int bsg_nifti_info(const char *name, int dims[3], int *datatype);

/* Volume 1 as doubles, x fastest (dim[1]*dim[2]*dim[3] values = n). Reads
 * UINT8, INT16, INT32, FLOAT32 and FLOAT64; big-endian files are swapped. As
 * in read_data.cpp:204-226, 342-370, 503-516 the scaling slope is not
 * applied; the data start at vox_offset (352 for the reference's files). */
int bsg_nifti_read(const char *name, double *out, size_t n);

/* write_nifti_file (nifti1_read_write.cpp:198-322) for one FLOAT64 volume:
 * sizeof_hdr 348, dim = {3, nx, ny, nz, 1}, bitpix 64, pixdim 2 mm,
 * vox_offset 352, xyzt_units mm, magic "n+1", a zero 4-byte extender, then
 * the data. A name ending in .gz is written gzip-compressed (the reference
 * writes .nii and runs `gzip -f`, mcmc.cpp:824-964). */
int bsg_nifti_write_f64(const char *name, int nx, int ny, int nz, const double *data);

/* Design file (read_data.cpp:262-316): whitespace-separated tokens; the first
 * ncovar + 2 are column names (subject id, ncovar covariate names, image
 * file), then one row of ncovar + 2 tokens per subject. Call once with NULL
 * outputs to get *nsubs, then with:
 *   covar   float[ncovar][nsubs] (atof of each covariate token)
 *   names   ncovar * 64 chars, covariate column names (NUL-terminated)
 *   images  nsubs * 512 chars, "./images/" + image token (read_data.cpp:304-306)
 * A token count that is not a whole number of rows is an error (the
 * reference silently truncates). */
int bsg_design_read(const char *path, int ncovar, int *nsubs, float *covar, char *names, char *images);

/* The BinCar program (main.cu:60-432 with the sampler driver of mcmc.cpp):
 * same arguments, inputs under ./images, seed.dat, console report and output
 * files (spatCoef_*, standCoef_*, prb_*, empir_prb_*, total_*, bWM, Qhat.dat,
 * Qhat2.dat, DIC.dat). Returns the process exit status. The executable
 * bsglmm_amd/BinCar is a main() that calls it. */
int bsg_bincar_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif
#endif
