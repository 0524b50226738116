/*
 * oracle.h -- CPU restatement of gpu_cross's FP64 cross-section path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against; it is never linked into, loaded by or called from the
 * product library (libgx.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * What it restates (reference files, relative to the reference checkout):
 *   constants            defines.h:4-20, GPUManager.cpp:28-31
 *   partition function   exomol_functions.cpp:50-73
 *   InitializeConstants  OpenMPManager.cpp:57-67
 *   line intensity       OpenMPManager.cpp:151-157 (Doppler), :196-202 (Voigt)
 *   Lorentz T-scaling    OpenMPManager.cpp:204
 *   Doppler width        OpenMPManager.cpp:146,191 ("openmp" convention),
 *                        cross_kernal.cu:667,938 ("gpu" convention)
 *   cutoff window        OpenMPManager.cpp:160-161,192,205-206 (round()),
 *                        cross_kernal.cu:686-687,954-957 (per-bin test)
 *   Voigt function       profiles.cpp:9-42 (Humlicek 1982 W4)
 *   Voigt accumulation   OpenMPManager.cpp:183-224
 *   Doppler accumulation OpenMPManager.cpp:138-181
 *   thread reduction     OpenMPManager.cpp:118-132
 *   chunk range          BaseProfile.cpp:55-57
 *   Gaussian profile     cross_kernal.cu:471-496 (intensity), :567-625 (accumulation),
 *                        GPUManager.cpp:24-47 (constants; ln2 taken from defines.h:4)
 *   Voigt without gamma  cross_kernal.cu:731-756 (intensity), :860-922 (accumulation),
 *                        :61-147 (Zaghloul & Ali Alg. 916), :183-186 (3-point Gauss-Hermite)
 *
 * Parity pins (tests/test_oracle.py): every function here is checked against
 * the golden fixtures written by the reference's own OpenMPManager (built
 * from its sources by oracle/Makefile into oracle/_ref), and against closed
 * forms (two-level partition sum, Voigt asymptotes and area, Doppler
 * telescoping sum, window brute force).
 */
#ifndef GX_ORACLE_H
#define GX_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* defines.h:4-20 -- literal values, reproduced bit for bit */
#define ORC_LN2 0.69314718056
#define ORC_ISQRTPI 0.564189584
#define ORC_PI 3.14159265359
#define ORC_PLANCK 6.6260693e-27
#define ORC_VELLGT 2.99792458e+10
#define ORC_BOLTZ 1.380658e-16
#define ORC_NA 6.0221412927e23
#define ORC_DOPPLER_CUTOFF 100.0
#define ORC_SQRTLN2PI 0.46971863935
#define ORC_SQRTLN2 0.83255461115

enum { ORC_GAUSSIAN = 0, ORC_DOPPLER = 1, ORC_VOIGT = 2 };   /* Input.h:9-15 numbering */
enum { ORC_CONV_OPENMP = 0, ORC_CONV_GPU = 1 };

typedef struct orc_consts {
    int profile;            /* ORC_DOPPLER or ORC_VOIGT */
    int convention;         /* ORC_CONV_OPENMP (CPU manager) or ORC_CONV_GPU */
    double temperature;     /* K */
    double partition;       /* Q(T) */
    double dfreq;           /* bin width, cm^-1 */
    double mean_mass;       /* amu */
    double pressure;        /* atm */
    double ref_temp;        /* K (500 for ExoMol decks, VoigtProfile.cpp:14) */
    double cutoff;          /* window half-width, cm^-1 */
    double doppler_widths;  /* >0: Doppler per-line cutoff = k*alpha_i + dfreq */
    double half_width;      /* Gaussian HWHM, cm^-1 (InitializeConstants' half_width) */
} orc_consts;

/* Humlicek W4 Re w(x+iy), profiles.cpp:9-42 */
double orc_humlic(double x, double y);

/* Q(T) = sum_s g_s exp(-beta E_s), exomol_functions.cpp:50-73 */
double orc_partition(size_t n, const double *energy, const int *gtot, double temperature);

/* beta = h c / (k T), OpenMPManager.cpp:66 */
double orc_beta(double temperature);
/* 1/(8 pi c) with the literal PI (openmp) or acos(-1) (gpu), OpenMPManager.cpp:65 / GPUManager.cpp:31,39 */
double orc_cmcoef(int convention);
/* Doppler width per unit wavenumber, OpenMPManager.cpp:191 / cross_kernal.cu:938 */
double orc_dpwcoef(double temperature, double mean_mass, int convention);
/* reference rule for the cutoff: Voigt min(25P,100) (OpenMPManager.cpp:192), Doppler 100 (defines.h:20) */
double orc_reference_cutoff(int profile, double pressure);

/* S_ij(T), OpenMPManager.cpp:151-157 / :196-202 (and the gpu-convention
 * quirks of cross_kernal.cu:645-646,774-777) */
double orc_line_intensity(const orc_consts *c, double elow, double nu, double aif, int gns);

/* gamma_L(T) = gamma * (T_ref/T)^n, OpenMPManager.cpp:204 */
double orc_gamma_t(const orc_consts *c, double gamma, double nexp);

/* Integer bin window [ib, ie) of one line in a chunk [start_idx, start_idx+points),
 * OpenMPManager.cpp:205-206 (openmp) or the per-bin |nu-f|<=C test of
 * cross_kernal.cu:954-957 (gpu; needs freq).  Returns ie-ib (may be <= 0). */
long orc_window(const orc_consts *c, double nu, double cutoff_i, const double *freq, int npoints_total,
                int points, int start_idx, int *ib, int *ie);

/* Per-line cutoff actually applied (cutoff, or doppler_widths*alpha + dfreq). */
double orc_line_cutoff(const orc_consts *c, double nu);

/* ExecuteCrossSection + TransferResults of the CPU manager for one chunk:
 * out[j] += sum over lines, with the reference's nthreads contiguous line
 * slices, private per-thread grids and ordered thread sum.  Runs the slices
 * on nthreads POSIX threads.  gamma/nexp may be NULL for Doppler.
 * Returns 0 on success.  freq has npoints_total entries. */
int orc_cross_section(const orc_consts *c, size_t nlines, const double *elow, const double *nu,
                      const double *aif, const int *gns, const double *gamma, const double *nexp,
                      const double *freq, int npoints_total, int points, int start_idx,
                      int nthreads, double *out);

/* Chunk bin range of BaseProfile.cpp:55-57 for a chunk whose min/max nu are given. */
void orc_chunk_range(double min_nu, double max_nu, double hw, double start_nu, double dfreq,
                     int npoints, int *ib, int *ie);

/* Count of line-bin evaluations sum_i (ie_i - ib_i)+ for the chunk. */
long long orc_count_evals(const orc_consts *c, size_t nlines, const double *nu, const double *freq,
                          int npoints_total, int points, int start_idx);

/* erfcx(y) = exp(y^2) erfc(y), y >= 0: exp of the exactly split square times libm erfc
 * below 26, Laplace's continued fraction (Abramowitz & Stegun 7.1.14) above. */
double orc_erfcx(double y);
/* Zaghloul & Ali (2011, ACM TOMS 38, Alg. 916) Voigt V(x,y) as cross_kernal.cu:61-147 states
 * it: the sigma series truncated to n = 1..4 around n0 = ceil(x/a). */
double orc_voigt916(double x, double y, double a);
/* the 3-point Gauss-Hermite expression of cross_kernal.cu:183-186, as written */
double orc_gh3(double x, double y, double xxyy);
/* S of the CUDA manager's Gaussian and no-gamma Voigt steps (cross_kernal.cu:471-496, 731-756):
 * cmcoef with acos(-1), no ISQRTPI, nu == 0 gives 0 */
double orc_line_intensity_plain(const orc_consts *c, double elow, double nu, double aif, int gns);

/* element-wise orc_humlic / libm erf over arrays (test convenience) */
void orc_voigt916_vec(size_t n, const double *x, const double *y, double a, double *out);
void orc_erfcx_vec(size_t n, const double *x, double *out);
void orc_humlic_vec(size_t n, const double *x, const double *y, double *out);
void orc_erf_vec(size_t n, const double *x, double *out);

#ifdef __cplusplus
}
#endif
#endif
