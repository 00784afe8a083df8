/*
 * hos_mkic.c -- writes a reference-format input pair (initpars.h5 and
 * initdata.{id}.h5) holding a seeded JONSWAP sea state, the role of the
 * reference's python/CreateSpectra.py + python/write_hdf5.py:5-40.
 *
 * initpars.h5 keys and types follow the reference's own fixture
 * (Csource/2dpar/initpars.h5): Nx, Ny, T, Tramp, rampflg, runsubid, saveflg,
 * KP as int64 scalars; Lx, Ly, g, dtsave, h as float64 scalars.
 * initdata.{id}.h5 holds /eta0, /phi0, /bat0 as float64 [Nx][Ny]; bat0 = 0.
 *
 *   hos_mkic Nx Ny Lx Ly g depth kp gamma eps seed T dtsave [runsubid] [out_dir]
 *
 * depth <= 0 means deep water (h is then written as 0).  Physical fields are
 * formed from the spectral state by the GPU library's inverse transform.
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/hos_gpu.h"
#include "../../include/hos_ic.h"

static void put_i64(hid_t f, const char *name, long long v)
{
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t d = H5Dcreate2(f, name, H5T_STD_I64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v);
    H5Dclose(d);
    H5Sclose(sp);
}

static void put_f64(hid_t f, const char *name, double v)
{
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t d = H5Dcreate2(f, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v);
    H5Dclose(d);
    H5Sclose(sp);
}

static void put_field(hid_t f, const char *name, const double *padded, int Nx, int Ny, double *tmp)
{
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) tmp[(size_t)Ny * i + j] = padded ? padded[(size_t)(Ny + 2) * i + j] : 0.0;
    hsize_t dim[2] = {(hsize_t)Nx, (hsize_t)Ny};
    hid_t sp = H5Screate_simple(2, dim, NULL);
    hid_t d = H5Dcreate2(f, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp);
    H5Dclose(d);
    H5Sclose(sp);
}

int main(int argc, char **argv)
{
    if (argc < 13) {
        fprintf(stderr, "usage: %s Nx Ny Lx Ly g depth kp gamma eps seed T dtsave [runsubid] [out_dir]\n", argv[0]);
        return 2;
    }
    hos_jonswap j;
    memset(&j, 0, sizeof j);
    j.Nx = atoi(argv[1]);
    j.Ny = atoi(argv[2]);
    j.Lx = atof(argv[3]);
    j.Ly = atof(argv[4]);
    j.g = atof(argv[5]);
    j.depth = atof(argv[6]);
    j.kp = atof(argv[7]);
    j.gamma = atof(argv[8]);
    j.eps = atof(argv[9]);
    j.seed = strtoull(argv[10], NULL, 10);
    j.kcut = 4 * j.kp;
    const double T = atof(argv[11]), dtsave = atof(argv[12]);
    const int id = argc > 13 ? atoi(argv[13]) : 1;
    if (argc > 14 && chdir(argv[14]) != 0) {
        perror(argv[14]);
        return 1;
    }
    const int Nx = j.Nx, Ny = j.Ny;
    const size_t fs = (size_t)Nx * (Ny / 2 + 1);
    double *u = calloc(4 * fs, sizeof(double));
    double *eta = calloc((size_t)Nx * (Ny + 2), sizeof(double)), *phi = calloc((size_t)Nx * (Ny + 2), sizeof(double));
    double *tmp = calloc((size_t)Nx * Ny, sizeof(double));
    if (!u || !eta || !phi || !tmp) return 1;
    if (hos_ic_jonswap(&j, u) != 0) {
        fprintf(stderr, "hos_mkic: invalid spectrum parameters\n");
        return 1;
    }
    hos_params p;
    hos_params_default(&p);
    p.Nx = Nx;
    p.Ny = Ny;
    p.Lx = j.Lx;
    p.Ly = j.Ly;
    hos_ctx *c = NULL;
    if (hos_create(&p, 1, &c) != HOS_OK || hos_ifft_2d(c, u, eta) != HOS_OK ||
        hos_ifft_2d(c, u + 2 * fs, phi) != HOS_OK) {
        fprintf(stderr, "hos_mkic: %s\n", hos_last_error());
        return 1;
    }
    hos_destroy(c);

    hid_t f = H5Fcreate("initpars.h5", H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    put_i64(f, "Nx", Nx);
    put_i64(f, "Ny", Ny);
    put_f64(f, "Lx", j.Lx);
    put_f64(f, "Ly", j.Ly);
    put_f64(f, "g", j.g);
    put_f64(f, "h", j.depth > 0 ? j.depth : 0.0);
    put_i64(f, "KP", (long long)j.kp);
    put_i64(f, "T", (long long)T);
    put_f64(f, "dtsave", dtsave);
    put_i64(f, "saveflg", 1);
    put_i64(f, "rampflg", 0);
    put_i64(f, "Tramp", 1);
    put_i64(f, "runsubid", id);
    H5Fclose(f);
    char name[64];
    snprintf(name, sizeof name, "initdata.%d.h5", id);
    f = H5Fcreate(name, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    put_field(f, "eta0", eta, Nx, Ny, tmp);
    put_field(f, "phi0", phi, Nx, Ny, tmp);
    put_field(f, "bat0", NULL, Nx, Ny, tmp);
    H5Fclose(f);
    printf("wrote initpars.h5 and %s (%dx%d)\n", name, Nx, Ny);
    free(u);
    free(eta);
    free(phi);
    free(tmp);
    return 0;
}
