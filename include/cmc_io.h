/* Initial-condition input of the Henon hot path: the reference's HDF5 cluster file.
 *
 * Replaces cmc_read_hdf5_file (src/libs/fitslib.c:209-447) plus the parts of
 * load_fits_file_data (src/cmc/cmc_fits.c:63-146) and the input block of
 * parser (src/cmc/cmc_io.c:2037-2072) that the hot path consumes, so a cluster
 * file written for the reference (COSMIC's InitialCMCTable.write, e.g.
 * ci/input.hdf5) starts a run here unchanged.
 *
 * File layout read (pandas "fixed" HDFStore, as the reference reads it):
 *   CLUS_OBJ_DATA/block0_values     float64 [NOBJ+2][8]
 *       columns id, k, m, Reff, r, vr, vt, binind; row 0 is the central mass,
 *       rows 1..NOBJ the objects sorted by r, row NOBJ+1 the outer sentinel
 *   attributes of that dataset      NOBJ, NBINARY (int64); RVIR, RTID, MCLUS, Z
 *       (read converted to double, as the reference's H5Aread does)
 *   CLUS_BINARY_DATA/block0_values  float64 [NBINARY+1][11]: not read here.
 *
 * Units and conventions follow load_fits_file_data: masses are scaled by NOBJ
 * (file masses are fractions of the cluster mass), the central mass is row 0's
 * scaled mass and row 0 of m is then zero, r[0] = 1e-20 and r[NOBJ+1] = 1e10
 * (ZERO and SF_INFINITY, include/cmc/cmc.h:58-60).  A binary is one object of
 * its total mass for the hot path; its internal state is outside it.
 *
 * Built as libcmcio.so against the HDF5 C library; nothing else in this
 * repository depends on it. */
#ifndef CMC_IO_H
#define CMC_IO_H

#include <stdint.h>
#include "cmcg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CMC_IO_OK        0
#define CMC_IO_E_OPEN   -1   /* file missing or not HDF5 (the reference exits(1)) */
#define CMC_IO_E_LAYOUT -2   /* dataset or attribute missing, wrong rank/columns */
#define CMC_IO_E_SIZE   -3   /* NOBJ disagrees with the rows, or the caller's n */
#define CMC_IO_E_READ   -4   /* HDF5 read failed */
#define CMC_IO_E_ARG    -5   /* null pointer */
#define CMC_IO_E_DATA   -6   /* a row breaks the layout (r not ascending, m < 0) */

typedef struct cmc_hdf5_header {
    int64_t nobj;      /* NOBJ: objects (singles + binaries), rows 1..nobj   */
    int64_t nbinary;   /* NBINARY                                            */
    int64_t rows;      /* rows of CLUS_OBJ_DATA/block0_values (nobj + 2)    */
    int64_t cols;      /* its columns (8)                                    */
    double rvir;       /* RVIR, pc: units.l (cmc_utils.c:1168)              */
    double rtid;       /* RTID: R_MAX, the initial orbit_r (cmc_io.c:2071)   */
    double mclus;      /* MCLUS, Msun: units.m (cmc_utils.c:1167)            */
    double z;          /* Z: METALLICITY (cmc_io.c:2072)                     */
    double cenma_m;    /* central mass, row 0's m * nobj (cmc_fits.c:134)   */
} cmc_hdf5_header;

/* Header only: the attributes and the table's shape. */
int cmc_hdf5_read_header(const char *path, cmc_hdf5_header *hdr);

/* Objects into out (n must equal hdr->nobj; arrays of n+2 entries): id, m (scaled),
 * r, vr, vt filled for slots 0..n+1 with the sentinel conventions above; other
 * arrays of out are left alone.  k, reff and binind, when not NULL, receive the
 * stellar type, radius and binary index columns for slots 0..n+1 as stored. */
int cmc_hdf5_read_objects(const char *path, const cmc_hdf5_header *hdr, cmcg_host_soa *out,
                          int32_t *k, double *reff, int64_t *binind);

const char *cmc_io_error_string(int code);

#ifdef __cplusplus
}
#endif

#endif
