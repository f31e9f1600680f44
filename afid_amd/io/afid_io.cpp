// afid_io.cpp -- continuation, statistics and slab files (include/afid_io.h) over the public C ABI
// of libafid_hip and the HDF5 C library.  Host code only.
//
// Layout followed: WriteFlowField.F90:17-45 (file names, master scalars),
// HdfRoutines.F90:25-96 (1-element scalar datasets, NATIVE_DOUBLE / 32-bit
// NATIVE_INTEGER), :388-477 (dataset 'var', Fortran dims (nx, nym, nzm), i.e.
// C dims {nzm, nym, nx}), ReadFlowField.F90:29-45 (master read), :160-170.
#include <hdf5.h>
#include <sys/stat.h>

#include <cmath>
#include <new>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/afid_io.h"

namespace {

struct H5Silence {  // keep HDF5's error stack off stderr; failures come back as codes
  H5E_auto2_t fn = nullptr;
  void *data = nullptr;
  H5Silence() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~H5Silence() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

std::string join(const char *dir, const char *name) {
  std::string d = dir && *dir ? dir : ".";
  if (d.back() != '/') d += '/';
  return d + name;
}

// Writes (creates or replaces) dataset `dset` in an open file.
int write_dataset(hid_t f, const char *dset, hid_t type, const void *buf, int rank, const hsize_t *dims) {
  if (H5Lexists(f, dset, H5P_DEFAULT) > 0 && H5Ldelete(f, dset, H5P_DEFAULT) < 0) return AFID_ERR_IO;
  hid_t sp = H5Screate_simple(rank, dims, nullptr);
  if (sp < 0) return AFID_ERR_IO;
  hid_t ds = H5Dcreate2(f, dset, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = AFID_OK;
  if (ds < 0 || H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) rc = AFID_ERR_IO;
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  return rc;
}

// Reads dataset `dset` whose dims must equal `dims`.
int read_dataset(hid_t f, const char *dset, hid_t type, void *buf, int rank, const hsize_t *dims) {
  hid_t ds = H5Dopen2(f, dset, H5P_DEFAULT);
  if (ds < 0) return AFID_ERR_IO;
  int rc = AFID_OK;
  hid_t sp = H5Dget_space(ds);
  hsize_t got[8] = {0};
  if (sp < 0 || H5Sget_simple_extent_ndims(sp) != rank || H5Sget_simple_extent_dims(sp, got, nullptr) != rank)
    rc = AFID_ERR_IO;
  for (int q = 0; rc == AFID_OK && q < rank; q++)
    if (got[q] != dims[q]) rc = AFID_ERR_IO;   // dataset of another shape
  if (rc == AFID_OK && H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) rc = AFID_ERR_IO;
  if (sp >= 0) H5Sclose(sp);
  H5Dclose(ds);
  return rc;
}

int write_field_file(const std::string &path, const char *dset, const double *host, int n1, int n2, int n3,
                     bool create) {
  hid_t f = create ? H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                   : H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) return AFID_ERR_IO;
  const hsize_t dims[3] = {(hsize_t)n3, (hsize_t)n2, (hsize_t)n1};
  int rc = write_dataset(f, dset, H5T_NATIVE_DOUBLE, host, 3, dims);
  if (H5Fclose(f) < 0 && rc == AFID_OK) rc = AFID_ERR_IO;
  return rc;
}

int read_field_file(const std::string &path, const char *dset, double *host, int n1, int n2, int n3) {
  hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return AFID_ERR_IO;
  const hsize_t dims[3] = {(hsize_t)n3, (hsize_t)n2, (hsize_t)n1};
  int rc = read_dataset(f, dset, H5T_NATIVE_DOUBLE, host, 3, dims);
  H5Fclose(f);
  return rc;
}

struct Dims {
  int nx, nxm, nym, nzm;
};

int dims_of(afid_ctx *ctx, Dims *d) {
  int64_t out[6];
  int rc = afid_get_dims(ctx, out);
  if (rc) return rc;
  d->nx = (int)out[0];
  d->nxm = (int)out[1];
  d->nym = (int)out[2];
  d->nzm = (int)out[3];
  return AFID_OK;
}

// The fields of a continuation: reference files first, then the state file.
struct FieldFile {
  const char *file;
  const char *dset;
  int field;
};
const FieldFile kReferenceFields[] = {{"continua_temp.h5", "var", AFID_TEMP},
                                      {"continua_vx.h5", "var", AFID_VX},
                                      {"continua_vy.h5", "var", AFID_VY},
                                      {"continua_vz.h5", "var", AFID_VZ}};
const FieldFile kStateFields[] = {{"continua_state.h5", "pr", AFID_PR},
                                  {"continua_state.h5", "rux", AFID_RUX},
                                  {"continua_state.h5", "ruy", AFID_RUY},
                                  {"continua_state.h5", "ruz", AFID_RUZ},
                                  {"continua_state.h5", "rutemp", AFID_RUTEMP}};

}  // namespace

extern "C" {

int afid_h5_write_field(const char *path, const char *dset, const double *host, int n1, int n2, int n3) {
  if (!path || !dset || !host || n1 < 1 || n2 < 1 || n3 < 1) return AFID_ERR_ARG;
  H5Silence quiet;
  return write_field_file(path, dset, host, n1, n2, n3, true);
}

int afid_h5_read_field(const char *path, const char *dset, double *host, int n1, int n2, int n3) {
  if (!path || !dset || !host || n1 < 1 || n2 < 1 || n3 < 1) return AFID_ERR_ARG;
  H5Silence quiet;
  return read_field_file(path, dset, host, n1, n2, n3);
}

int afid_write_continuation(afid_ctx *ctx, const char *dir) {
  if (!ctx) return AFID_ERR_ARG;
  H5Silence quiet;
  Dims d;
  int rc = dims_of(ctx, &d);
  if (rc) return rc;
  std::vector<double> a((size_t)d.nx * d.nym * d.nzm);
  for (const FieldFile &ff : kReferenceFields) {
    if ((rc = afid_download(ctx, ff.field, a.data(), d.nx, d.nym, 1, 1))) return rc;
    if ((rc = write_field_file(join(dir, ff.file), ff.dset, a.data(), d.nx, d.nym, d.nzm, true))) return rc;
  }
  bool first = true;
  for (const FieldFile &ff : kStateFields) {
    if ((rc = afid_download(ctx, ff.field, a.data(), d.nx, d.nym, 1, 1))) return rc;
    if ((rc = write_field_file(join(dir, ff.file), ff.dset, a.data(), d.nx, d.nym, d.nzm, first))) return rc;
    first = false;
  }
  {  // stale row rhs(1,:,:), Fortran shape (nym, nzm)
    std::vector<double> plane((size_t)d.nym * d.nzm);
    if ((rc = afid_get_stale(ctx, plane.data()))) return rc;
    hid_t f = H5Fopen(join(dir, "continua_state.h5").c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) return AFID_ERR_IO;
    const hsize_t dims[2] = {(hsize_t)d.nzm, (hsize_t)d.nym};
    rc = write_dataset(f, "rhs1", H5T_NATIVE_DOUBLE, plane.data(), 2, dims);
    H5Fclose(f);
    if (rc) return rc;
  }
  // continua_master.h5: the scalars of WriteFlowField.F90:26-45
  double time = 0.0, ylen = 0.0, zlen = 0.0, str3 = 0.0;
  if ((rc = afid_get_time(ctx, &time))) return rc;
  double v[1];
  if ((rc = afid_get_metric(ctx, "ylen", v, 1))) return rc;
  ylen = v[0];
  if ((rc = afid_get_metric(ctx, "zlen", v, 1))) return rc;
  zlen = v[0];
  if ((rc = afid_get_metric(ctx, "str3", v, 1))) return rc;
  str3 = v[0];
  if ((rc = afid_get_metric(ctx, "istr3", v, 1))) return rc;
  const int istr3 = (int)v[0];
  hid_t f = H5Fcreate(join(dir, "continua_master.h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return AFID_ERR_IO;
  const hsize_t one[1] = {1};
  const int nx = d.nx, ny = d.nym + 1, nz = d.nzm + 1;
  rc = write_dataset(f, "nx", H5T_NATIVE_INT, &nx, 1, one);
  if (!rc) rc = write_dataset(f, "ny", H5T_NATIVE_INT, &ny, 1, one);
  if (!rc) rc = write_dataset(f, "nz", H5T_NATIVE_INT, &nz, 1, one);
  if (!rc) rc = write_dataset(f, "ylen", H5T_NATIVE_DOUBLE, &ylen, 1, one);
  if (!rc) rc = write_dataset(f, "zlen", H5T_NATIVE_DOUBLE, &zlen, 1, one);
  if (!rc) rc = write_dataset(f, "time", H5T_NATIVE_DOUBLE, &time, 1, one);
  if (!rc) rc = write_dataset(f, "istr3", H5T_NATIVE_INT, &istr3, 1, one);
  if (!rc) rc = write_dataset(f, "str3", H5T_NATIVE_DOUBLE, &str3, 1, one);
  if (H5Fclose(f) < 0 && !rc) rc = AFID_ERR_IO;
  return rc;
}

int afid_write_stats(afid_ctx *ctx, const char *dir, int readstats, int disscal) {
  if (!ctx) return AFID_ERR_ARG;
  H5Silence quiet;
  Dims d;
  int rc = dims_of(ctx, &d);
  if (rc) return rc;
  const int nxm = d.nxm;
  std::vector<double> st((size_t)11 * nxm);
  int nsamples = 0;
  if ((rc = afid_get_stats(ctx, st.data(), &nsamples))) return rc;
  const char *names[11] = {"vx_mean", "vy_mean", "vz_mean", "temp_mean", "vx_rms", "vy_rms",
                           "vz_rms", "temp_rms", "tempvx_mean", "disste", "dissth"};
  const int nq = disscal ? 11 : 9;
  const std::string path = join(dir, "stafield_master.h5");
  const hsize_t one[1] = {1}, dn[1] = {(hsize_t)nxm};
  hid_t f = -1;
  if (readstats) f = H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
  if (f >= 0) {   // add what the file holds
    int old = 0;
    if ((rc = read_dataset(f, "averaging_time", H5T_NATIVE_INT, &old, 1, one))) {
      H5Fclose(f);
      return rc;
    }
    nsamples += old;
    std::vector<double> prev(nxm);
    for (int q = 0; q < nq; q++) {
      if ((rc = read_dataset(f, names[q], H5T_NATIVE_DOUBLE, prev.data(), 1, dn))) {
        H5Fclose(f);
        return rc;
      }
      for (int k = 0; k < nxm; k++) st[(size_t)q * nxm + k] += prev[k];
    }
  } else {
    f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return AFID_ERR_IO;
  }
  for (int q = 0; q < nq && !rc; q++) rc = write_dataset(f, names[q], H5T_NATIVE_DOUBLE, &st[(size_t)q * nxm], 1, dn);
  std::vector<double> xm(nxm);
  double ray = 0.0, pra = 0.0;
  if (!rc) rc = write_dataset(f, "averaging_time", H5T_NATIVE_INT, &nsamples, 1, one);
  if (!rc) rc = afid_get_metric(ctx, "xm", xm.data(), nxm);
  if (!rc) rc = write_dataset(f, "X_cordin", H5T_NATIVE_DOUBLE, xm.data(), 1, dn);
  if (!rc) rc = afid_get_metric(ctx, "ray", &ray, 1);
  if (!rc) rc = write_dataset(f, "Rayleigh Number", H5T_NATIVE_DOUBLE, &ray, 1, one);
  if (!rc) rc = afid_get_metric(ctx, "pra", &pra, 1);
  if (!rc) rc = write_dataset(f, "Prandtl Number", H5T_NATIVE_DOUBLE, &pra, 1, one);
  if (H5Fclose(f) < 0 && !rc) rc = AFID_ERR_IO;
  return rc;
}

int afid_slab_levels(const double *xm, int nxm, const double *xslab8, int *kslab9) {
  if (!xm || !xslab8 || !kslab9 || nxm < 2) return AFID_ERR_ARG;
  for (int m = 0; m < 9; m++) kslab9[m] = 2;
  for (int k = 2; k <= nxm; k++)
    for (int j = 2; j <= 9; j++)
      if (xm[k - 1] > xslab8[j - 2] && xm[k - 2] < xslab8[j - 2]) kslab9[j - 1] = k;
  return AFID_OK;
}

int afid_write_slabs(afid_ctx *ctx, const char *dir, const int *kslab9, double time, double tout) {
  if (!ctx || !kslab9 || tout <= 0) return AFID_ERR_ARG;
  H5Silence quiet;
  Dims d;
  int rc = dims_of(ctx, &d);
  if (rc) return rc;
  const size_t plane = (size_t)d.nym * d.nzm;
  mkdir(dir, 0755);   // the reference expects ./stst3 to exist; an existing one is kept
  std::vector<double> buf(9 * 4 * plane);
  if ((rc = afid_get_slabs(ctx, kslab9, 9, buf.data()))) return rc;
  const long itime = std::lround(time * (1.0 / tout));
  const char *qn[4] = {"vx", "vy", "vz", "temp"};
  const hsize_t dims[2] = {(hsize_t)d.nzm, (hsize_t)d.nym}, one[1] = {1};
  for (int m = 0; m < 9; m++)
    for (int q = 0; q < 4; q++) {
      char name[64];
      std::snprintf(name, sizeof name, "slab%d%s_%08ld.h5", m + 1, qn[q], itime);
      hid_t f = H5Fcreate(join(dir, name).c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
      if (f < 0) return AFID_ERR_IO;
      rc = write_dataset(f, "var", H5T_NATIVE_DOUBLE, &buf[(m * 4 + q) * plane], 2, dims);
      if (!rc) rc = write_dataset(f, "time", H5T_NATIVE_DOUBLE, &time, 1, one);
      if (H5Fclose(f) < 0 && !rc) rc = AFID_ERR_IO;
      if (rc) return rc;
    }
  return AFID_OK;
}

int afid_read_continuation(afid_ctx *ctx, const char *dir) {
  if (!ctx) return AFID_ERR_ARG;
  H5Silence quiet;
  Dims d;
  int rc = dims_of(ctx, &d);
  if (rc) return rc;
  // master: grid check (ReadFlowField.F90:56-58) and time
  hid_t f = H5Fopen(join(dir, "continua_master.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return AFID_ERR_IO;
  const hsize_t one[1] = {1};
  int nxo = 0, nyo = 0, nzo = 0, istro3 = 0;
  double time = 0.0, stro3 = 0.0;
  rc = read_dataset(f, "nx", H5T_NATIVE_INT, &nxo, 1, one);
  if (!rc) rc = read_dataset(f, "ny", H5T_NATIVE_INT, &nyo, 1, one);
  if (!rc) rc = read_dataset(f, "nz", H5T_NATIVE_INT, &nzo, 1, one);
  if (!rc) rc = read_dataset(f, "time", H5T_NATIVE_DOUBLE, &time, 1, one);
  if (!rc) rc = read_dataset(f, "istr3", H5T_NATIVE_INT, &istro3, 1, one);
  if (!rc) rc = read_dataset(f, "str3", H5T_NATIVE_DOUBLE, &stro3, 1, one);
  H5Fclose(f);
  if (rc) return rc;
  double v[1];
  if ((rc = afid_get_metric(ctx, "str3", v, 1))) return rc;
  const double str3 = v[0];
  if ((rc = afid_get_metric(ctx, "istr3", v, 1))) return rc;
  std::vector<double> plane((size_t)d.nym * d.nzm, 0.0);
  if (nxo != d.nx || nyo != d.nym + 1 || nzo != d.nzm + 1 || istro3 != (int)v[0] || std::fabs(stro3 - str3) > 1e-8) {
    // another grid: read each field on its own grid and interpolate
    // (ReadFlowField.F90:64-158); pr, the histories and the stale row start
    // from zero, as the reference restarts
    // sizes from the file: refuse implausible ones before allocating
    if (nxo < 2 || nyo < 3 || nzo < 3 || nxo > 65536 || nyo > 65536 || nzo > 65536 ||
        (double)nxo * nyo * nzo > 6.9e10)
      return AFID_ERR_IO;
    std::vector<double> o;
    try {
      o.resize((size_t)nxo * (nyo - 1) * (nzo - 1));
    } catch (const std::bad_alloc &) {
      return AFID_ERR_NOMEM;
    }
    for (const FieldFile &ff : kReferenceFields) {
      if ((rc = read_field_file(join(dir, ff.file), ff.dset, o.data(), nxo, nyo - 1, nzo - 1))) return rc;
      if ((rc = afid_interp_field(ctx, ff.field, o.data(), nxo, nyo, nzo, istro3, stro3))) return rc;
    }
    std::vector<double> z((size_t)d.nx * d.nym * d.nzm, 0.0);
    for (const FieldFile &ff : kStateFields)
      if ((rc = afid_upload(ctx, ff.field, z.data(), d.nx, d.nym, 1, 1))) return rc;
    if ((rc = afid_set_stale(ctx, plane.data()))) return rc;
    return afid_set_time(ctx, time);
  }
  std::vector<double> a((size_t)d.nx * d.nym * d.nzm);
  for (const FieldFile &ff : kReferenceFields) {
    if ((rc = read_field_file(join(dir, ff.file), ff.dset, a.data(), d.nx, d.nym, d.nzm))) return rc;
    if ((rc = afid_upload(ctx, ff.field, a.data(), d.nx, d.nym, 1, 1))) return rc;
  }
  const std::string state = join(dir, "continua_state.h5");
  hid_t fs = H5Fopen(state.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fs >= 0) {
    H5Fclose(fs);
    for (const FieldFile &ff : kStateFields) {
      if ((rc = read_field_file(state, ff.dset, a.data(), d.nx, d.nym, d.nzm))) return rc;
      if ((rc = afid_upload(ctx, ff.field, a.data(), d.nx, d.nym, 1, 1))) return rc;
    }
    hid_t f2 = H5Fopen(state.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    const hsize_t dims[2] = {(hsize_t)d.nzm, (hsize_t)d.nym};
    rc = read_dataset(f2, "rhs1", H5T_NATIVE_DOUBLE, plane.data(), 2, dims);
    H5Fclose(f2);
    if (rc) return rc;
  } else {  // the reference restarts from pr = 0 and zeroed histories
    std::fill(a.begin(), a.end(), 0.0);
    for (const FieldFile &ff : kStateFields)
      if ((rc = afid_upload(ctx, ff.field, a.data(), d.nx, d.nym, 1, 1))) return rc;
  }
  if ((rc = afid_set_stale(ctx, plane.data()))) return rc;
  return afid_set_time(ctx, time);
}

}  // extern "C"
