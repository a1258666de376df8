/*
 * textio.c -- the reference's text output format on the host.
 *
 * GDLOutputCylinder (output.h:280-384) writes one "%e\n" per value, interior
 * cells only, the second index fastest.  The same format is what the
 * reference's IDL scripts read with fltarr + readf (density.pro:27-83).
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>

#include "../../include/cylmhd.h"
#include "textio.h"

int cyl_mkdir(const char *dir)
{
    if (mkdir(dir, S_IRUSR | S_IWUSR | S_IXUSR | S_IRGRP | S_IWGRP | S_IXGRP | S_IROTH | S_IXOTH) != 0 &&
        errno != EEXIST)
        return CYLMHD_EIO;
    return CYLMHD_OK;
}

int cyl_write_text(const char *dir, const char *name, const double *v, size_t n)
{
    char path[4096];
    if (snprintf(path, sizeof path, "%s/%s", dir, name) >= (int)sizeof path) return CYLMHD_EINVAL;
    FILE *fo = fopen(path, "w");
    if (!fo) return CYLMHD_EIO;
    /* a large stdio buffer: a 512 x 512 plane is 262144 short lines */
    setvbuf(fo, NULL, _IOFBF, 1 << 20);
    int bad = 0;
    for (size_t q = 0; q < n; q++)
        if (fprintf(fo, "%e\n", v[q]) < 0) bad = 1;
    if (fclose(fo) != 0) bad = 1;
    return bad ? CYLMHD_EIO : CYLMHD_OK;
}

int cyl_write_grid_text(const char *dir, const char *rname, const char *phiname, const char *zname,
                        int nx, int ny, int nz, const double *x, const double *y, const double *z)
{
    /* output.h:288-314: x[1..ixmax1-1], y[1..jymax1-1], z[1..kzmax1-1] */
    int rc;
    if ((rc = cyl_write_text(dir, rname, x + 1, (size_t)(nx - 2)))) return rc;
    if ((rc = cyl_write_text(dir, phiname, y + 1, (size_t)(ny - 2)))) return rc;
    return cyl_write_text(dir, zname, z + 1, (size_t)(nz - 2));
}

/* GDLOutputCylinder (output.h:280-384), on host arrays in the reference
 * layout V[9][nx][ny][nz]. */
int cylmhd_gdl_output(const char *dir, int nout, int nx, int ny, int nz,
                      const double *x, const double *y, const double *z, const double *V)
{
    if (!dir || !x || !y || !z || !V || nx < 3 || ny < 3 || nz < 3) return CYLMHD_EINVAL;
    int rc;
    if ((rc = cyl_mkdir(dir))) return rc;
    if ((rc = cyl_write_grid_text(dir, "r.dat", "phi.dat", "z.dat", nx, ny, nz, x, y, z))) return rc;
    const size_t N = (size_t)nx * ny * nz, ni = (size_t)nx - 2, nj = (size_t)ny - 2, nk = (size_t)nz - 2;
    size_t cap = ni * nj > ni * nk ? ni * nj : ni * nk;
    double *buf = (double *)malloc(cap * sizeof(double));
    if (!buf) return CYLMHD_ENOMEM;
    char name[128];
#define AT(m, i, j, k) V[(size_t)(m) * N + ((size_t)(i) * ny + (size_t)(j)) * nz + (size_t)(k)]
    /* k = 1 density, pressure, magnetic energy (output.h:318-362) */
    static const char *what[3] = {"density", "pressure", "benergy"};
    for (int w = 0; w < 3 && !rc; w++) {
        size_t q = 0;
        for (int i = 1; i < nx - 1; i++)
            for (int j = 1; j < ny - 1; j++) {
                const int k = 1;
                buf[q++] = w == 0 ? AT(0, i, j, k)
                         : w == 1 ? AT(7, i, j, k)
                                  : 0.5 * (AT(4, i, j, k) * AT(4, i, j, k) + AT(5, i, j, k) * AT(5, i, j, k)
                                           + AT(6, i, j, k) * AT(6, i, j, k));
            }
        snprintf(name, sizeof name, "2d-%s-rphi-%d.dat", what[w], nout);
        rc = cyl_write_text(dir, name, buf, q);
    }
    /* j = 0 density, k outer, i inner (output.h:366-380) */
    if (!rc) {
        size_t q = 0;
        for (int k = 1; k < nz - 1; k++)
            for (int i = 1; i < nx - 1; i++) buf[q++] = AT(0, i, 0, k);
        snprintf(name, sizeof name, "2d-density-rz-%d.dat", nout);
        rc = cyl_write_text(dir, name, buf, q);
    }
#undef AT
    free(buf);
    return rc;
}
