/* Writes the reference's mesh file layout (`domain/mesh_utils.py:155-194` _write_xdmf, written
 * there with h5py): datasets /mesh/points (float64 N x 3), /mesh/cells (int64 C x 4) and
 * /mesh/cell_data (int32 C) in an HDF5 file, with the HDF5 C library itself, in four storage
 * variants, so that the reader of dsv.meshio is checked against files the library wrote:
 *   0  default properties (h5py's defaults: earliest format, contiguous datasets)
 *   1  chunked datasets (no filters)
 *   2  latest file format (version-3 superblock, version-2 object headers, link messages)
 *   3  big-endian file datatypes
 * Test infrastructure: built and run by tests/golden/make_h5_fixtures.py.
 * Usage: h5_writer VARIANT OUT.h5 points.bin cells.bin data.bin NPTS NCELLS */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>

static void *slurp(const char *path, size_t n) {
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); exit(1); }
    void *p = malloc(n);
    if (fread(p, 1, n, f) != n) { fprintf(stderr, "short read %s\n", path); exit(1); }
    fclose(f);
    return p;
}

static void write(hid_t grp, const char *name, hid_t ftype, hid_t mtype, int rank,
                  const hsize_t *dims, const void *buf, int chunked) {
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    if (chunked) {
        hsize_t ch[2] = {dims[0] > 7 ? 7 : dims[0], rank > 1 ? dims[1] : 1};
        H5Pset_chunk(dcpl, rank, ch);
    }
    hid_t ds = H5Dcreate2(grp, name, ftype, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (ds < 0 || H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) exit(2);
    H5Dclose(ds);
    H5Pclose(dcpl);
    H5Sclose(sp);
}

int main(int argc, char **argv) {
    if (argc != 8) { fprintf(stderr, "usage\n"); return 1; }
    const int variant = atoi(argv[1]);
    const size_t npts = (size_t)atol(argv[6]), ncells = (size_t)atol(argv[7]);
    double *pts = slurp(argv[3], npts * 3 * sizeof(double));
    long long *cells = slurp(argv[4], ncells * 4 * sizeof(long long));
    int *data = slurp(argv[5], ncells * sizeof(int));
    hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
    if (variant == 2) H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST);
    hid_t f = H5Fcreate(argv[2], H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
    hid_t g = H5Gcreate2(f, "mesh", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const int big = variant == 3, chunked = variant == 1;
    hsize_t dp[2] = {npts, 3}, dc[2] = {ncells, 4}, dd[1] = {ncells};
    write(g, "points", big ? H5T_IEEE_F64BE : H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, dp, pts, chunked);
    write(g, "cells", big ? H5T_STD_I64BE : H5T_STD_I64LE, H5T_NATIVE_LLONG, 2, dc, cells, chunked);
    write(g, "cell_data", big ? H5T_STD_I32BE : H5T_STD_I32LE, H5T_NATIVE_INT, 1, dd, data, chunked);
    H5Gclose(g);
    H5Fclose(f);
    H5Pclose(fapl);
    free(pts); free(cells); free(data);
    return 0;
}
