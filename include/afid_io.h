/*
 * afid_io.h -- continuation, statistics and slab files of a cell
 * (libafid_io.so, HDF5).
 *
 * The reference writes one file per prognostic field, continua_<name>.h5 with
 * a dataset 'var' of Fortran shape (nx, nym, nzm), and continua_master.h5
 * with the scalars nx, ny, nz, ylen, zlen, time, istr3, str3
 * (WriteFlowField.F90:17-45, HdfRoutines.F90:388-477).  These entry points
 * write and read the same layout.  They also write continua_state.h5: pr,
 * the histories rux..rutemp and the stale row rhs(1,:,:).  The reference
 * restarts without them, from pr = 0 and zeroed histories
 * (ReadFlowField.F90:160-170).  With them a restart continues bit for bit; a
 * directory without that file restarts as the reference does.
 *
 * When continua_master.h5 names another grid (ReadFlowField.F90:64-66),
 * the reader interpolates each field onto this grid with afid_interp_field
 * (ReadFlowField.F90:67-158) and starts pr, the histories and the stale row
 * from zero, as the reference does.
 *
 * Return codes as in afid.h; AFID_ERR_IO for a missing or malformed file,
 * AFID_ERR_GRID for an old grid that cannot be interpolated from (see
 * afid_interp_field); the fields may then be partly set.
 */
#ifndef AFID_IO_H
#define AFID_IO_H
#include "afid.h"
#ifdef __cplusplus
extern "C" {
#endif

/* WriteFlowField.F90:11-45 into directory dir (time from afid_get_time). */
int afid_write_continuation(afid_ctx *ctx, const char *dir);
/* ReadFlowField.F90:29-170 from directory dir: fields (interpolated from
 * another grid), time (afid_set_time), and the state file when present and
 * the grid is the same. */
int afid_read_continuation(afid_ctx *ctx, const char *dir);

/* WriteStats (StatRoutines.F90:49-137): stafield_master.h5 in dir with the
 * accumulated vx_mean .. tempvx_mean (and disste, dissth when disscal != 0),
 * averaging_time = number of samples, X_cordin = xm(1:nxm), and the Rayleigh
 * and Prandtl numbers.  With readstats != 0 and an existing file, the arrays
 * and averaging_time already in it are added first
 * (StatReadReduceWrite.F90:21-28, StatRoutines.F90:95-107). */
int afid_write_stats(afid_ctx *ctx, const char *dir, int readstats, int disscal);
/* InitializeSlabDump (SlabDumpRoutines.F90:40-70): kslab[0] = 2 and, for each
 * xslab[q] (q = 0..7, the reference's xslab(2..9)), the last k in 2..nxm with
 * xm(k-1) < xslab < xm(k), 2 where there is none.  xm is 1-based: xm[0] is
 * xm(1).  Host only. */
int afid_slab_levels(const double *xm, int nxm, const double *xslab8, int *kslab9);
/* SlabDumper + DumpSingleSlab (SlabDumpRoutines.F90:12-133): for the nine
 * levels kslab writes slab<m><vx|vy|vz|temp>_<itime>.h5 into dir (m = 1..9,
 * itime = nint(time/tout) as 8 digits), each with dataset 'var' of Fortran
 * shape (nym, nzm) and the scalar 'time'.  dir is created if missing. */
int afid_write_slabs(afid_ctx *ctx, const char *dir, const int *kslab9, double time, double tout);

/* One FP64 dataset of Fortran shape (n1, n2, n3), column-major host array
 * (HdfWriteRealHalo3D / HdfReadRealHalo3D, HdfRoutines.F90:299-477, serial). */
int afid_h5_write_field(const char *path, const char *dset, const double *host, int n1, int n2, int n3);
int afid_h5_read_field(const char *path, const char *dset, double *host, int n1, int n2, int n3);

#ifdef __cplusplus
}
#endif
#endif
