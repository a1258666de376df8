/*
 * textio.h -- text-slice helpers shared by the host writers (internal).
 */
#ifndef CYLMHD_TEXTIO_H
#define CYLMHD_TEXTIO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mkdir that accepts an existing directory; CYLMHD_OK or CYLMHD_EIO */
int cyl_mkdir(const char *dir);
/* <dir>/<name>: n values, "%e\n" each (the reference's text format,
 * output.h:288-380); CYLMHD_OK or CYLMHD_EIO */
int cyl_write_text(const char *dir, const char *name, const double *v, size_t n);
/* the three interior coordinate lists r.dat, phi.dat, z.dat (output.h:286-314)
 * under the given names */
int cyl_write_grid_text(const char *dir, const char *rname, const char *phiname, const char *zname,
                        int nx, int ny, int nz, const double *x, const double *y, const double *z);

#ifdef __cplusplus
}
#endif
#endif
