/* C ABI of the model's on-disk records: the history/restart record and the ECMWF
 * forcing records.  Host-side, no GPU.  Both are Fortran sequential unformatted files
 * written big-endian (the model is built with -convert big_endian,
 * reprobus-simulation.sh:133; the forcing with gfortran -fconvert=big-endian,
 * reprobus-extract-ecmwf.sh:146): each record is framed by its byte length as a 4-byte
 * big-endian integer before and after it.  All arrays are Fortran order (lon fastest);
 * grid sizes are arguments (the model: nlon = 180, nlat = 91, niv = 137).
 *
 * Return codes are those of rbchem.h (RB_OK, RB_EINVAL; RB_EIO for a file that cannot be
 * opened, is short, or has inconsistent record framing); rb_io_last_error() describes
 * the last failure of the calling thread. */
#ifndef RBIO_H
#define RBIO_H
#include <stdint.h>

#include "rbchem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RB_EIO (-6)

/* History/restart record (write F:587-588 to unit 78 'history', read F:231-233 from
 * unit 90; layout also MODEL_history_convert_to_netcdf.py:113-129): one record holding
 * namexp (6 characters), ian, imois, ijour, iheure, imin (int32), then pj1 (nlon x nlat),
 * uj1, vj1, alt, tj1 (field each), qj1 (43 fields), surfarea (field), hc (15 fields),
 * all real*8, field = nlon x nlat x niv.  The record must stay below 2^31 bytes. */
typedef struct {
    char namexp[7];          /* 6 characters + NUL */
    int32_t ian, imois, ijour, iheure, imin;
} rb_history_header;

int rb_history_write(const char *path, int32_t nlon, int32_t nlat, int32_t niv,
                     const rb_history_header *h, const double *pj1, const double *uj1,
                     const double *vj1, const double *alt, const double *tj1, const double *qj1,
                     const double *surfarea, const double *hc);
int rb_history_read(const char *path, int32_t nlon, int32_t nlat, int32_t niv,
                    rb_history_header *h, double *pj1, double *uj1, double *vj1, double *alt,
                    double *tj1, double *qj1, double *surfarea, double *hc);

/* ECMWF forcing file ecmwf_YYYYMMDD (read F:4713-4935, written by
 * reprobus-extract-ecmwf.sh:108-146): per analysis time, a record of 18 int32 (the
 * GRIB header: idate(10..13) = year, month, day, hour; idate(16) = step hours) and a
 * record of the surface pressure (nlon x nlat real*8), then for each level 1..niv five
 * (header, field) record pairs: T, W, H (specific humidity), U, V (nlon x nlat each).
 *
 * rb_ecmwf_write_time appends one analysis time (append = 0 truncates the file first);
 * idate_ps is the surface-pressure header, idate_lev (niv x 5 x 18, level-major) the
 * headers of the level fields.  rb_ecmwf_read_time reads analysis time `itime` (0-based,
 * the earlier ones skipped as F:4765-4773 does), returning the surface header in idate
 * (18) and the fields; H is floored at 1e-30*18/28.97 as the model does (F:4793-4797). */
int rb_ecmwf_write_time(const char *path, int32_t append, int32_t nlon, int32_t nlat,
                        int32_t niv, const int32_t *idate_ps, const double *ps,
                        const int32_t *idate_lev, const double *t, const double *w,
                        const double *h, const double *u, const double *v);
int rb_ecmwf_read_time(const char *path, int32_t itime, int32_t nlon, int32_t nlat, int32_t niv,
                       int32_t *idate, double *ps, double *t, double *w, double *h, double *u,
                       double *v);

const char *rb_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
