/*
 * gx_output.h -- C ABI of libgx.so's output writers: the step after the
 * cross-section path.
 *
 *   gx_xsec_text*   the reference's output, one "%12.6f %13.8E\n" line (frequency,
 *                   cross-section) per bin (BaseProfile::OutputProfile,
 *                   BaseProfile.cpp:65-74).
 *   gx_opacity_*    an opacity-table file for the batched (T,P) tables, which the
 *                   reference (one point per process) has no format for.
 *
 * Opacity-table file, little-endian:
 *   offset  0  char[8]  "GXOPAC01"
 *           8  u32      header bytes (128)
 *          12  u32      version (1)
 *          16  u64      npoints            bins per grid
 *          24  u64      npts               (T, P) points
 *          32  f64      start_nu, end_nu   grid: f_i = start_nu + i (end_nu - start_nu) / npoints
 *          48  f64      dfreq
 *          56  i32      profile (GX_DOPPLER, GX_VOIGT), i32 precision (GX_PREC_*)
 *          64  f64      mean_mass, ref_temp, cutoff
 *          88  u64      data_offset        start of the grids, a multiple of 4096
 *          96  zeros
 *         128  f64[npts][3]               (T, P, Q) of each point
 *  data_offset f64[npts][npoints]         the grids, point after point
 * Conventions as gx.h (gx_status results, gx_last_error()).
 */
#ifndef GX_OUTPUT_H
#define GX_OUTPUT_H

#include <stddef.h>
#include <stdint.h>

#include "gx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Format n "%12.6f %13.8E\n" lines into out (cap bytes); *written = bytes needed. */
int gx_xsec_text(const double *freq, const double *intens, int n, char *out, size_t cap, size_t *written);
/* Write them to a file (path NULL: standard output, as the reference prints them). */
int gx_xsec_text_write(const char *path, const double *freq, const double *intens, int n);

typedef struct gx_opacity_info {
    uint64_t npoints;
    uint64_t npts;
    double start_nu;
    double end_nu;
    double dfreq;
    int32_t profile;
    int32_t precision;
    double mean_mass;
    double ref_temp;
    double cutoff;
    uint64_t data_offset;   /* filled by gx_opacity_write / gx_opacity_read_info */
} gx_opacity_info;

/* Write a table: info (data_offset ignored), npts points and npts * npoints grid values. */
int gx_opacity_write(const char *path, const gx_opacity_info *info, const gx_point *pts, const double *grids);
/* Read a table file's header. */
int gx_opacity_read_info(const char *path, gx_opacity_info *info);
/* Read its points (npts) and grids (npts * npoints values). */
int gx_opacity_read(const char *path, gx_point *pts, size_t npts, double *grids, size_t nvalues);

#ifdef __cplusplus
}
#endif

#endif /* GX_OUTPUT_H */
