/*
 * gx_tips.h -- total internal partition sums (TIPS-2011) for HITRAN line lists.
 *
 * Replaces the reference's Fortran TIPS routine bd_tips_2011_ (declared in
 * HITRAN_func.h:5, defined in HITRAN_files/BD_TIPS_2011_v1p0.for:6-347, its
 * interpolation AtoB at :5830-5922), which HITRANStateReader calls for Q(T) and
 * Q(296 K) (HITRANStateReader.cpp:62, 95).  The tables are the TIPS-2011 data
 * (Fischer et al., JQSRT 82 (2003) 401; Gamache's 2011 update): this library
 * ships none of them and reads them from the caller's copy of that Fortran
 * source file, so Q(T) comes from the same numbers the reference links in.
 *
 * Reading: each QofT/xgj DATA constant is a default-REAL (single precision)
 * literal stored into a DOUBLE PRECISION array, so its value is the nearest
 * float, widened (as ifort without -fpconstant compiles it).  Q(T) is the
 * 4-point Lagrange interpolation on the 119-point temperature grid, 3-point at
 * either end, in the routine's own operation order.
 *
 * Host code; plain pointers and sizes; status codes as in gx.h.
 */
#ifndef GX_TIPS_H
#define GX_TIPS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gx_tips gx_tips;

/* Parse a TIPS-2011 Fortran source file (BD_TIPS_2011_v1p0.for). */
int gx_tips_load(const char *path, gx_tips **out);
/* The same from text already in memory. */
int gx_tips_load_text(const char *text, size_t n, gx_tips **out);
void gx_tips_free(gx_tips *tips);

/* bd_tips_2011_(MOL, Temp, ISO, gi, QT) (BD_TIPS_2011_v1p0.for:6-347): HITRAN molecule
 * number mol, isotopologue index iso (1-based, HITRAN order), temperature T in K.  Writes
 * the state-independent degeneracy gi and Q(T).  Molecules 34 (O) and 39 (CH3OH) give
 * gi = 0, Q = 0 as the routine does.  T outside [70, 3000] K: Q = -1 (the routine's value)
 * and GX_INVALID_ARG.  An unknown molecule or isotopologue: GX_INVALID_ARG (the routine
 * leaves QT unset or reads past its table). */
int gx_tips_q(const gx_tips *tips, int mol, int iso, double temperature, double *gi, double *q);

/* Isotopologues tabulated for molecule mol (0 when none). */
int gx_tips_isotopes(const gx_tips *tips, int mol);

/* Lagrange interpolation of the routine's AtoB (BD_TIPS_2011_v1p0.for:5830-5922) on any
 * table: n points a[] (increasing), values b[], at x; 3-point within the first two
 * intervals and at the last node, 4-point elsewhere.  x above a[n-1] gives *out = 0 and
 * GX_INVALID_ARG (the routine then leaves bb unset). */
int gx_tips_lagrange(const double *a, const double *b, int n, double x, double *out);

#ifdef __cplusplus
}
#endif
#endif /* GX_TIPS_H */
