/*
 * gx_hitran.h -- C ABI of libgx.so's HITRAN .par reader: the other reader the
 * reference's Voigt driver takes its lines from.
 *
 * Replaces, with the same results:
 *   GetFieldStart / GetFieldLength / GetField   HITRANStateReader.cpp:3-40
 *                    (fixed fields 2,1,12,10,10,5,5,10,4,8,15,15,15,15,6,12,1,7,7 characters;
 *                    a field is copied into a zeroed 13-byte array, cut at 12, and read
 *                    with atof / atoi)
 *   the first record's molecule and isotope     HITRANStateReader.cpp:42-58, 77-86
 *   ReadNextState (6 args)                      HITRANStateReader.cpp:136-250
 *                    (one fgets(buffer, 1024) piece per record: nu = field 2, E'' = 7,
 *                    A = 4, gamma = (gamma_air x + gamma_self (1 - x)) P from fields 5, 6,
 *                    n = 8, g = 17; a record with A == 0 takes A from its intensity
 *                    (field 3) with Q(296 K) and g = 1)
 *   the read loop's filter and storage          VoigtProfile.cpp:61-75
 *                    (`continue` below lo, `break` above hi; g stored into an int array)
 *
 * The reference's reader keeps one 1024-byte buffer for the whole file, so a record
 * shorter than the fields it reads takes the missing bytes from earlier, longer records
 * (its last fgets left them there); this reader does the same, with the buffer zeroed at
 * the start of each text (the reference leaves it uninitialised).  Q(T) and Q(296 K) come
 * from the caller: the reference computes them with HITRAN's TIPS Fortran routine
 * (bd_tips_2011_), which is not part of this library.
 *
 * Records are parsed in parallel on the host (a .par file holds at most a few million
 * 160-byte records).  Conventions as gx.h: gx_status results, gx_last_error(), borrowed
 * host pointers.
 */
#ifndef GX_HITRAN_H
#define GX_HITRAN_H

#include <stddef.h>

#include "gx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gx_hitran_opts {
    double pressure;        /* atm: multiplies gamma (HITRANStateReader m_pressure) */
    double mixture;         /* air fraction x of gamma_air x + gamma_self (1 - x) */
    double ref_partition;   /* Q(296 K): A of records whose A field is 0 */
    double lo, hi;          /* keep nu >= lo, stop at the first nu > hi */
    int nthreads;           /* <= 0: all cores */
} gx_hitran_opts;

/* Read .par text.  With cap == 0 (arrays may be NULL) only *nlines is computed; else up to
 * cap kept lines are written in read order (E'', nu, A, g, gamma, n: TransferVectors'
 * arrays; gamma / n may be NULL).  *nrecords = records read (ReadNextState calls that
 * returned true, up to and including the one above hi); mol_id / iso (may be NULL) = atoi
 * of the first record's fields 0 and 1. */
int gx_hitran_read(const char *text, size_t len, const gx_hitran_opts *opts, size_t cap, double *elow, double *nu,
                   double *aif, int *gns, double *gamma, double *nexp, size_t *nlines, size_t *nrecords, int *mol_id,
                   int *iso);

/* HITRAN 160-character records for n lines, fields as HITRAN's format lays them out (the
 * writer of the synthetic inputs): nu %12.6f, S %10.3E, A %10.3E, gamma_air %5.4f,
 * gamma_self %5.3f, E'' %10.4f, n %4.2f, delta %8.6f, the quanta, error and reference fields
 * blank, g' %7.1f, g'' %7.1f.  out must hold 161 n bytes. */
int gx_format_hitran(size_t n, int mol_id, int iso, const double *nu, const double *intens, const double *aif,
                     const double *gamma_air, const double *gamma_self, const double *elow, const double *nexp,
                     const double *gup, const double *glow, char *out, size_t cap, size_t *written);

#ifdef __cplusplus
}
#endif

#endif /* GX_HITRAN_H */
