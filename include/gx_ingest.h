/*
 * gx_ingest.h -- C ABI of libgx.so's ExoMol reader: the step before the
 * cross-section path, which turns ExoMol .states / .trans / .broad text into
 * the line arrays BaseManager::TransferVectors takes (BaseManager.h:24).
 *
 * Replaces, with the same results:
 *   initialize_states        exomol_functions.cpp:8-48    (.states lines -> E, g, J)
 *   compute_partition        exomol_functions.cpp:50-73   (Q(T), also
 *                            ExomolStateReader::ComputePartition, ExomolStateReader.cpp:172-178)
 *   InitializeBroadener      ExomolStateReader.cpp:39-87  (.broad lines -> key map)
 *   ReadNextState (6 args)   ExomolStateReader.cpp:130-167 (one .trans record -> nu, g, E'', A,
 *                            gamma, n, with GetGammaN, ExomolStateReader.cpp:7-37)
 *   the read loop's filter   VoigtProfile.cpp:61-68 (`continue` below lo, `break` above hi)
 *
 * The .trans records are tokenised and converted on the GPU: the caller hands over text
 * (a whole file, or consecutive pieces of it) and the kept lines collect in HBM.  The
 * reference reads its records with `stream >> id_f >> id_i >> aif`; the same tokens give
 * the same numbers here, and reading ends where the stream's would (a token the stream
 * cannot read, the first line above hi).  Conventions as gx.h: gx_status results,
 * gx_last_error(), borrowed host pointers.
 */
#ifndef GX_INGEST_H
#define GX_INGEST_H

#include <stddef.h>

#include "gx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Q(T) = sum_s g_s exp(-beta E_s), beta = h c / (k T), summed in state order with the
 * constants of exomol_functions.cpp:55-60 (replaces compute_partition,
 * exomol_functions.cpp:50-73). */
int gx_partition(const double *energy, const int *gtot, size_t nstates, double temperature, double *q);

/* Number of state records in .states text: one per fgets(buf, 1024) piece, so a line
 * longer than 1022 characters counts once per 1023-character piece
 * (exomol_functions.cpp:15-18). */
int gx_states_count(const char *text, size_t len, size_t *nstates);

/* Parse .states text into caller arrays of gx_states_count() entries: per piece
 * id = strtol(.., 0) - 1, E = strtod, g = strtol(.., 0), J = strtol(.., 0)
 * (exomol_functions.cpp:28-36).  Any of id/J may be NULL.  nthreads <= 0: all cores. */
int gx_states_parse(const char *text, size_t len, size_t nstates, int nthreads, int *id, double *energy,
                    int *gtot, int *J);

typedef struct gx_ingest gx_ingest;

/* Reader over one molecule's states on a device: the states (positional, as
 * exomol.states[id - 1], ExomolStateReader.cpp:156-160) are copied to HBM. */
int gx_ingest_create(int device, size_t nstates, const double *energy, const int *gtot, const int *J,
                     gx_ingest **out);
void gx_ingest_destroy(gx_ingest *ing);

/* .broad text and the pressure its gammas are multiplied by (InitializeBroadener and
 * GetGammaN, ExomolStateReader.cpp:7-87, with the reference's key strings: "a1" + Ji + Jf,
 * then "a1" + Ji + Jf + "a0" + Ji + "-1" since key.clear() keeps the text).  Without it
 * (or with len == 0) gamma and n are never assigned, as when the reference finds no
 * broadener file. */
int gx_ingest_set_broadening(gx_ingest *ing, const char *text, size_t len, double pressure);

/* Keep lines with nu >= lo; stop reading the current file at the first line with nu > hi
 * (VoigtProfile.cpp:65-68; lo = start - C, hi = end + C). */
int gx_ingest_set_range(gx_ingest *ing, double lo, double hi);

/* Start a new .trans file (the reader reopens its stream, VoigtProfile.cpp:57). */
int gx_ingest_begin_file(gx_ingest *ing);

/* The next piece of the current file's text.  Pieces may split anywhere; the reader
 * keeps an unfinished record for the next call.  eof != 0 marks the file's last piece. */
int gx_ingest_feed(gx_ingest *ing, const char *text, size_t len, int eof);

/* Lines kept so far (all files since creation or the last gx_ingest_clear), and records
 * read (ReadNextState calls that returned true). */
int gx_ingest_count(gx_ingest *ing, size_t *nlines, size_t *nrecords);

/* Copy the kept lines out in read order: TransferVectors' arrays (E'', nu, A, g, gamma, n).
 * gamma / n may be NULL. */
int gx_ingest_download(gx_ingest *ing, size_t n, double *elow, double *nu, double *aif, int *gns, double *gamma,
                       double *nexp);

/* Forget the kept lines (the reader's gamma/n carry and the range stay). */
int gx_ingest_clear(gx_ingest *ing);

/* Device time of the last feed's kernels (ms) and its bytes of text. */
int gx_ingest_last_feed(gx_ingest *ing, float *kernel_ms, size_t *text_bytes, size_t *records);

/* ExoMol .trans lines "%12d %12d %10.4E\n" for n transitions (the writer of the synthetic
 * inputs); out must hold 37 n bytes.  Returns the bytes written in *written. */
int gx_format_trans(size_t n, const int *id_f, const int *id_i, const double *aif, char *out, size_t cap,
                    size_t *written);

#ifdef __cplusplus
}
#endif

#endif /* GX_INGEST_H */
