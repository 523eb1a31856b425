/* Reader of the reference's HDF5 cluster file (include/cmc_io.h).
 *
 * Follows cmc_read_hdf5_file (src/libs/fitslib.c:209-447): the object table
 * CLUS_OBJ_DATA/block0_values and the scalar attributes NOBJ, NBINARY, RVIR,
 * RTID, MCLUS, Z of that dataset, every value read converted to the type the
 * reference asks of the library (doubles, int64 for the counts).  The mapping to
 * the hot path's arrays follows load_fits_file_data (src/cmc/cmc_fits.c:63-146).
 * Unlike the reference, every failure is returned as an error code instead of
 * ending the process, and the shapes are checked before anything is read. */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>

#include "cmc_io.h"

#define OBJ_PATH "CLUS_OBJ_DATA/block0_values"
#define OBJ_COLS 8
#define R_ZERO 1.0e-20        /* ZERO, include/cmc/cmc.h:60 */
#define R_INFINITY 1.0e10     /* SF_INFINITY, include/cmc/cmc.h:58 */

enum { C_ID = 0, C_K, C_M, C_REFF, C_R, C_VR, C_VT, C_BININD };

const char *cmc_io_error_string(int code)
{
    switch (code) {
    case CMC_IO_OK: return "ok";
    case CMC_IO_E_OPEN: return "cannot open the file as HDF5";
    case CMC_IO_E_LAYOUT: return "CLUS_OBJ_DATA/block0_values or one of its attributes is missing or malformed";
    case CMC_IO_E_SIZE: return "NOBJ does not match the table rows or the caller's n";
    case CMC_IO_E_READ: return "HDF5 read failed";
    case CMC_IO_E_ARG: return "null argument";
    case CMC_IO_E_DATA: return "object table breaks the layout (r not ascending or a negative mass)";
    default: return "unknown error";
    }
}

/* One scalar attribute of the object dataset, converted to mem_type. */
static int read_attr(hid_t dset, const char *name, hid_t mem_type, void *dst)
{
    if (H5Aexists(dset, name) <= 0) return CMC_IO_E_LAYOUT;
    hid_t a = H5Aopen(dset, name, H5P_DEFAULT);
    if (a < 0) return CMC_IO_E_LAYOUT;
    hid_t sp = H5Aget_space(a);
    hssize_t npts = sp >= 0 ? H5Sget_simple_extent_npoints(sp) : -1;
    int rc = CMC_IO_OK;
    if (npts != 1) rc = CMC_IO_E_LAYOUT;
    else if (H5Aread(a, mem_type, dst) < 0) rc = CMC_IO_E_READ;
    if (sp >= 0) H5Sclose(sp);
    H5Aclose(a);
    return rc;
}

/* Opens the file and the object dataset; fills hdr.  On success *file and *dset
 * are open and the caller closes them. */
static int open_objects(const char *path, cmc_hdf5_header *hdr, hid_t *file, hid_t *dset)
{
    *file = *dset = -1;
    if (!path || !hdr) return CMC_IO_E_ARG;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);     /* failures are reported as codes */
    FILE *probe = fopen(path, "rb");
    if (!probe) return CMC_IO_E_OPEN;
    fclose(probe);
    if (H5Fis_hdf5(path) <= 0) return CMC_IO_E_OPEN;
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return CMC_IO_E_OPEN;
    int rc = CMC_IO_OK;
    hid_t d = -1;
    if (H5Lexists(f, "CLUS_OBJ_DATA", H5P_DEFAULT) <= 0 || H5Lexists(f, OBJ_PATH, H5P_DEFAULT) <= 0) {
        rc = CMC_IO_E_LAYOUT;
        goto fail;
    }
    d = H5Dopen2(f, OBJ_PATH, H5P_DEFAULT);
    if (d < 0) { rc = CMC_IO_E_LAYOUT; goto fail; }
    {
        hid_t sp = H5Dget_space(d);
        hsize_t dims[2] = {0, 0};
        int nd = sp >= 0 ? H5Sget_simple_extent_ndims(sp) : -1;
        if (nd == 2) H5Sget_simple_extent_dims(sp, dims, NULL);
        if (sp >= 0) H5Sclose(sp);
        if (nd != 2 || dims[1] != OBJ_COLS) { rc = CMC_IO_E_LAYOUT; goto fail; }
        hdr->rows = (int64_t)dims[0];
        hdr->cols = (int64_t)dims[1];
    }
    long long nobj = 0, nbin = 0;
    if ((rc = read_attr(d, "NOBJ", H5T_NATIVE_LLONG, &nobj)) ||
        (rc = read_attr(d, "NBINARY", H5T_NATIVE_LLONG, &nbin)) ||
        (rc = read_attr(d, "RVIR", H5T_NATIVE_DOUBLE, &hdr->rvir)) ||
        (rc = read_attr(d, "RTID", H5T_NATIVE_DOUBLE, &hdr->rtid)) ||
        (rc = read_attr(d, "MCLUS", H5T_NATIVE_DOUBLE, &hdr->mclus)) ||
        (rc = read_attr(d, "Z", H5T_NATIVE_DOUBLE, &hdr->z)))
        goto fail;
    hdr->nobj = nobj;
    hdr->nbinary = nbin;
    /* load_fits_file_data reads rows 0..NOBJ+1 (cmc_fits.c:121-124). */
    if (nobj < 1 || hdr->rows != nobj + 2) { rc = CMC_IO_E_SIZE; goto fail; }
    *file = f;
    *dset = d;
    return CMC_IO_OK;
fail:
    if (d >= 0) H5Dclose(d);
    H5Fclose(f);
    return rc;
}

static int read_table(hid_t d, int64_t rows, double **out)
{
    double *t = (double *)malloc((size_t)rows * OBJ_COLS * sizeof(double));
    if (!t) return CMC_IO_E_READ;
    if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, t) < 0) {
        free(t);
        return CMC_IO_E_READ;
    }
    *out = t;
    return CMC_IO_OK;
}

int cmc_hdf5_read_header(const char *path, cmc_hdf5_header *hdr)
{
    hid_t f, d;
    int rc = open_objects(path, hdr, &f, &d);
    if (rc) return rc;
    /* Row 0 only (the central mass), through a one-row hyperslab. */
    double row0[OBJ_COLS];
    hsize_t start[2] = {0, 0}, count[2] = {1, OBJ_COLS};
    hid_t fsp = H5Dget_space(d);
    hid_t msp = H5Screate_simple(2, count, NULL);
    if (fsp < 0 || msp < 0 || H5Sselect_hyperslab(fsp, H5S_SELECT_SET, start, NULL, count, NULL) < 0 ||
        H5Dread(d, H5T_NATIVE_DOUBLE, msp, fsp, H5P_DEFAULT, row0) < 0)
        rc = CMC_IO_E_READ;
    else
        hdr->cenma_m = row0[C_M] * (double)hdr->nobj;
    if (msp >= 0) H5Sclose(msp);
    if (fsp >= 0) H5Sclose(fsp);
    H5Dclose(d);
    H5Fclose(f);
    return rc;
}

int cmc_hdf5_read_objects(const char *path, const cmc_hdf5_header *want, cmcg_host_soa *out,
                          int32_t *k, double *reff, int64_t *binind)
{
    if (!path || !want || !out || !out->r || !out->m || !out->vr || !out->vt || !out->id)
        return CMC_IO_E_ARG;
    cmc_hdf5_header hdr = {0};
    hid_t f, d;
    int rc = open_objects(path, &hdr, &f, &d);
    if (rc) return rc;
    double *t = NULL;
    if (hdr.nobj != want->nobj || out->n != hdr.nobj) rc = CMC_IO_E_SIZE;
    else rc = read_table(d, hdr.rows, &t);
    H5Dclose(d);
    H5Fclose(f);
    if (rc) return rc;

    const int64_t n = hdr.nobj;
    const double scale = (double)n;            /* clus.N_STAR, cmc_fits.c:99 */
    for (int64_t i = 1; i <= n; ++i) {
        const double *row = t + i * OBJ_COLS;
        if (!(row[C_M] >= 0.0) || (i > 1 && !(row[C_R] >= t[(i - 1) * OBJ_COLS + C_R]))) {
            free(t);
            return CMC_IO_E_DATA;
        }
    }
    for (int64_t i = 0; i <= n + 1; ++i) {
        const double *row = t + i * OBJ_COLS;
        out->id[i] = (int64_t)row[C_ID];
        out->m[i] = row[C_M] * scale;
        out->r[i] = row[C_R];
        out->vr[i] = row[C_VR];
        out->vt[i] = row[C_VT];
        if (k) k[i] = (int32_t)row[C_K];
        if (reff) reff[i] = row[C_REFF];
        if (binind) binind[i] = (int64_t)row[C_BININD];
    }
    /* Central mass leaves the star table; sentinels of the r array
     * (cmc_fits.c:127-136). */
    out->m[0] = 0.0;
    out->r[0] = R_ZERO;
    out->r[n + 1] = R_INFINITY;
    out->m[n + 1] = 0.0;
    free(t);
    return CMC_IO_OK;
}
