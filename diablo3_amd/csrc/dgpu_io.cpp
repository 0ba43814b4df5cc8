// dgpu_io.cpp — flow checkpoints (WriteHDF5 / ReadHDF5, for/phdf5.f90:918-1456)
// and statistics series (WriteStatH5_Y, WriteHDF5_real, for/phdf5.f90:7-278) in
// the reference's HDF5 layout, for one rank, through the serial HDF5 C API.
//
// Layout (what the reference's Fortran HDF5 calls produce; the Fortran API lists
// dimensions fastest first, the C API slowest first):
//   /               attributes Resolution int32[3] = (Nx, Ny, Nz), Date char[20]
//   /Timestep       attributes Time, Save_Flow_Time, Save_Stats_Time,
//                   Save_Movie_Time (f64), Time_Step (int32); datasets U, W, V,
//                   TH1..THn [, P], f64, C dims (Nz, Ny, Nx), chunks (Nz, 1, Nx):
//                   element [z][y-1][x] = field(x, z, y) (swapzy)
//   series groups   attribute SAMPLES int32; datasets "0001", "0002", ... (f64,
//                   Nyp values with attribute Time, or one value)
#include "dgpu_io.h"

#include <hdf5.h>
#include <sys/stat.h>

#include <cstdio>
#include <ctime>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// an HDF5 identifier closed by its own close function on scope exit
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid() = default;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Hid(const Hid&) = delete;
  Hid& operator=(const Hid&) = delete;
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
  operator hid_t() const { return id; }
};

struct Dims {
  int nx, ny, nz;
  size_t at(int x, int z, int y) const {  // physical (0:Nx+1, 0:Nzp+1, 0:Nyp+1)
    return (size_t)x + (size_t)(nx + 2) * ((size_t)z + (size_t)(nz + 2) * (size_t)y);
  }
  size_t field() const { return (size_t)(nx + 2) * (nz + 2) * (ny + 2); }
  size_t file() const { return (size_t)nx * ny * nz; }
};

bool dims_ok(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0; }

// swapzy (for/tools.f90:216-232): out(x+1, y, z+1) = in(x, z, y), x < Nx,
// y = 1..Nyp, z < Nzp; out in C order [z][y-1][x]
void swapzy(const Dims& d, const double* in, std::vector<double>& out) {
  out.assign(d.file(), 0.0);
  for (int z = 0; z < d.nz; ++z)
    for (int y = 1; y <= d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) out[((size_t)z * d.ny + (y - 1)) * d.nx + x] = in[d.at(x, z, y)];
}
// swapyz (for/tools.f90:235-251): the array zeroed, then the inverse map
void swapyz(const Dims& d, const std::vector<double>& in, double* out) {
  for (size_t i = 0; i < d.field(); ++i) out[i] = 0.0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 1; y <= d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) out[d.at(x, z, y)] = in[((size_t)z * d.ny + (y - 1)) * d.nx + x];
}

// time_string (for/tools.f90:155-211): 'DD-MON-YYYY HH:MM:SS', blanks of the
// day and of the time fields as '0'
std::string time_string() {
  static const char* mon[] = {"JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"};
  std::time_t now = std::time(nullptr);
  std::tm tmv{};
  localtime_r(&now, &tmv);
  char buf[32];
  std::snprintf(buf, sizeof buf, "%02d-%s-%4d %02d:%02d:%02d", tmv.tm_mday, mon[tmv.tm_mon], tmv.tm_year + 1900,
                tmv.tm_hour, tmv.tm_min, tmv.tm_sec);
  return std::string(buf, 20);
}

herr_t write_scalar_attr(hid_t obj, const char* name, hid_t ftype, hid_t mtype, const void* v) {
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  if (!sp.ok()) return -1;
  Hid a(H5Acreate2(obj, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return -1;
  return H5Awrite(a, mtype, v);
}
herr_t read_attr(hid_t obj, const char* name, hid_t mtype, void* v) {
  Hid a(H5Aopen_by_name(obj, ".", name, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return -1;
  return H5Aread(a, mtype, v);
}

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

bool file_exists(const char* fname) {
  struct stat st;
  return stat(fname, &st) == 0;
}

// the series group of WriteStatH5_Y / WriteHDF5_real: open (or create with
// SAMPLES = 0), bump SAMPLES, return the new sample number
int series_sample(hid_t f, const char* gname, hid_t& gid_out) {
  const std::string path = std::string("/") + gname;
  const htri_t ex = H5Lexists(f, path.c_str(), H5P_DEFAULT);
  if (ex < 0) return -1;
  int nsamp = 0;
  hid_t g;
  if (!ex) {
    g = H5Gcreate2(f, gname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return -1;
    if (write_scalar_attr(g, "SAMPLES", H5T_STD_I32LE, H5T_NATIVE_INT, &nsamp) < 0) { H5Gclose(g); return -1; }
  } else {
    g = H5Gopen2(f, path.c_str(), H5P_DEFAULT);
    if (g < 0) return -1;
    if (read_attr(g, "SAMPLES", H5T_NATIVE_INT, &nsamp) < 0) { H5Gclose(g); return -1; }
  }
  nsamp += 1;
  {
    Hid a(H5Aopen(g, "SAMPLES", H5P_DEFAULT), H5Aclose);
    if (!a.ok() || H5Awrite(a, H5T_NATIVE_INT, &nsamp) < 0) { H5Gclose(g); return -1; }
  }
  gid_out = g;
  return nsamp;
}

// open fname read-write, creating it first if absent (the reference's
// inquire / h5fcreate_f / h5fopen_f sequence)
hid_t open_series_file(const char* fname) {
  if (!file_exists(fname)) {
    hid_t f = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return -1;
    H5Fclose(f);
  }
  return H5Fopen(fname, H5F_ACC_RDWR, H5P_DEFAULT);
}

}  // namespace

extern "C" {

const char* dgpu_io_last_error(void) { return g_err.c_str(); }

void dgpu_g2gf(double* var, int32_t nx, int32_t ny, int32_t nz) {
  const Dims d{nx, ny, nz};
  for (int x = 0; x < nx; ++x)
    for (int z = 0; z < nz; ++z) {
      var[d.at(x, z, 1)] = var[d.at(x, z, 2)];            // rankY == 0
      var[d.at(x, z, ny + 1)] = var[d.at(x, z, ny)];      // rankY == NprocY - 1
      for (int y = 1; y <= ny; ++y) var[d.at(x, z, y)] = 0.5 * (var[d.at(x, z, y)] + var[d.at(x, z, y + 1)]);
    }
}

void dgpu_gf2g(double* var, int32_t nx, int32_t ny, int32_t nz) {
  const Dims d{nx, ny, nz};
  for (int x = 0; x < nx; ++x)
    for (int z = 0; z < nz; ++z) var[d.at(x, z, ny + 1)] = var[d.at(x, z, ny)];  // rankY == NprocY - 1
  for (int x = 0; x < nx; ++x)
    for (int z = 0; z < nz; ++z)
      for (int y = ny; y >= 1; --y) var[d.at(x, z, y)] = 2.0 * var[d.at(x, z, y)] - var[d.at(x, z, y + 1)];
  // zero mass flux at the walls, on every (x, z) of the array (var(:, :, 1), var(:, :, Nyp+1))
  for (int z = 0; z < nz + 2; ++z)
    for (int x = 0; x < nx + 2; ++x) {
      var[d.at(x, z, 1)] = -var[d.at(x, z, 2)];
      var[d.at(x, z, ny + 1)] = -var[d.at(x, z, ny)];
    }
}

int dgpu_h5_write_flow(const char* fname, int32_t nx, int32_t ny, int32_t nz, int32_t nth, const double* u1,
                       const double* u2f, const double* u3, const double* th, const double* p,
                       const dgpu_flow_times* t, const char* date20) {
  if (!fname || !u1 || !u2f || !u3 || !t || (nth > 0 && !th) || nth < 0 || nth > 9)
    return fail(1, "dgpu_h5_write_flow: null argument or bad n_th");
  if (!dims_ok(nx, ny, nz)) return fail(2, "dgpu_h5_write_flow: bad sizes");
  quiet();
  const Dims d{nx, ny, nz};
  Hid f(H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(3, std::string("cannot create ") + fname);
  {  // Resolution
    const hsize_t n3 = 3;
    Hid sp(H5Screate_simple(1, &n3, nullptr), H5Sclose);
    Hid a(H5Acreate2(f, "Resolution", H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    const int tint[3] = {nx, ny, nz};
    if (!a.ok() || H5Awrite(a, H5T_NATIVE_INT, tint) < 0) return fail(3, "attribute Resolution");
  }
  {  // Date: 20 one-character strings
    const std::string ds = date20 ? std::string(date20, 20) : time_string();
    const hsize_t n20 = 20;
    Hid sp(H5Screate_simple(1, &n20, nullptr), H5Sclose);
    Hid a(H5Acreate2(f, "Date", H5T_C_S1, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok() || H5Awrite(a, H5T_C_S1, ds.data()) < 0) return fail(3, "attribute Date");
  }
  Hid g(H5Gcreate2(f, "Timestep", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail(3, "group Timestep");
  if (write_scalar_attr(g, "Time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &t->time) < 0 ||
      write_scalar_attr(g, "Save_Flow_Time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &t->save_flow_time) < 0 ||
      write_scalar_attr(g, "Save_Stats_Time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &t->save_stats_time) < 0 ||
      write_scalar_attr(g, "Save_Movie_Time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &t->save_movie_time) < 0 ||
      write_scalar_attr(g, "Time_Step", H5T_STD_I32LE, H5T_NATIVE_INT, &t->time_step) < 0)
    return fail(3, "Timestep attributes");
  const hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
  const hsize_t chunk[3] = {(hsize_t)nz, 1, (hsize_t)nx};
  Hid space(H5Screate_simple(3, dims, nullptr), H5Sclose);
  Hid dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
  if (!space.ok() || !dcpl.ok() || H5Pset_chunk(dcpl, 3, chunk) < 0) return fail(3, "dataspace / chunking");
  std::vector<double> tmp;
  auto put = [&](const char* name, const double* field) -> int {
    swapzy(d, field, tmp);
    Hid ds(H5Dcreate2(g, name, H5T_IEEE_F64LE, space, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose);
    if (!ds.ok() || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp.data()) < 0)
      return fail(3, std::string("dataset ") + name);
    return 0;
  };
  int rc;
  if ((rc = put("U", u1)) || (rc = put("W", u2f)) || (rc = put("V", u3))) return rc;
  for (int n = 0; n < nth; ++n) {
    const std::string name = "TH" + std::to_string(n + 1);  // "TH"//char(ith + 45)
    if ((rc = put(name.c_str(), th + (size_t)n * d.field()))) return rc;
  }
  if (p && (rc = put("P", p))) return rc;
  return 0;
}

int dgpu_h5_read_flow(const char* fname, int32_t nx, int32_t ny, int32_t nz, int32_t nth, uint32_t th_mask,
                      double* u1, double* u2f, double* u3, double* th, double* p, int32_t* has_p,
                      dgpu_flow_times* t) {
  if (!fname || !u1 || !u2f || !u3 || !t || !has_p || (nth > 0 && !th) || nth < 0 || nth > 9)
    return fail(1, "dgpu_h5_read_flow: null argument or bad n_th");
  if (!dims_ok(nx, ny, nz)) return fail(2, "dgpu_h5_read_flow: bad sizes");
  quiet();
  const Dims d{nx, ny, nz};
  Hid f(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(3, std::string("cannot open ") + fname);
  int tint[3] = {0, 0, 0};
  if (read_attr(f, "Resolution", H5T_NATIVE_INT, tint) < 0) return fail(3, "attribute Resolution");
  if (tint[0] != nx || tint[1] != ny || tint[2] != nz)
    return fail(4, "file and program have different resolution: program " + std::to_string(nx) + " " +
                       std::to_string(ny) + " " + std::to_string(nz) + ", file " + std::to_string(tint[0]) + " " +
                       std::to_string(tint[1]) + " " + std::to_string(tint[2]));
  Hid g(H5Gopen2(f, "/Timestep", H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail(3, "group Timestep");
  if (read_attr(g, "Time", H5T_NATIVE_DOUBLE, &t->time) < 0 ||
      read_attr(g, "Save_Flow_Time", H5T_NATIVE_DOUBLE, &t->save_flow_time) < 0 ||
      read_attr(g, "Save_Stats_Time", H5T_NATIVE_DOUBLE, &t->save_stats_time) < 0 ||
      read_attr(g, "Save_Movie_Time", H5T_NATIVE_DOUBLE, &t->save_movie_time) < 0 ||
      read_attr(g, "Time_Step", H5T_NATIVE_INT, &t->time_step) < 0)
    return fail(3, "Timestep attributes");
  std::vector<double> tmp(d.file());
  auto get = [&](const char* name, double* field) -> int {
    Hid ds(H5Dopen2(g, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(3, std::string("dataset ") + name);
    Hid sp(H5Dget_space(ds), H5Sclose);
    hsize_t dims[3] = {0, 0, 0};
    if (!sp.ok() || H5Sget_simple_extent_ndims(sp) != 3 || H5Sget_simple_extent_dims(sp, dims, nullptr) < 0 ||
        dims[0] != (hsize_t)nz || dims[1] != (hsize_t)ny || dims[2] != (hsize_t)nx)
      return fail(4, std::string("dataset ") + name + " has the wrong shape");
    if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp.data()) < 0)
      return fail(3, std::string("reading ") + name);
    swapyz(d, tmp, field);
    return 0;
  };
  int rc;
  if ((rc = get("U", u1)) || (rc = get("W", u2f)) || (rc = get("V", u3))) return rc;
  for (int n = 0; n < nth; ++n) {
    if (!((th_mask >> n) & 1u)) continue;
    const std::string name = "TH" + std::to_string(n + 1);
    if ((rc = get(name.c_str(), th + (size_t)n * d.field()))) return rc;
  }
  const htri_t hp = H5Lexists(g, "P", H5P_DEFAULT);
  if (hp < 0) return fail(3, "probing P");
  *has_p = hp > 0 ? 1 : 0;
  if (*has_p) {
    if (!p) return fail(1, "dgpu_h5_read_flow: the file holds P but p is null");
    if ((rc = get("P", p))) return rc;
  }
  return 0;
}

namespace {
// one more sample "%04d" (C dims `dims`, f64) with its Time attribute in the
// series group gname, the group's SAMPLES counter advanced
int append_series(const char* fname, const char* gname, const double* data, int rank, const hsize_t* dims,
                  double time) {
  quiet();
  Hid f(open_series_file(fname), H5Fclose);
  if (!f.ok()) return fail(3, std::string("cannot open ") + fname);
  hid_t gid = -1;
  const int nsamp = series_sample(f, gname, gid);
  if (nsamp < 0) return fail(3, std::string("series group ") + gname);
  Hid g(gid, H5Gclose);
  char dname[32];
  std::snprintf(dname, sizeof dname, "%04d", nsamp);  // (1I0.4)
  Hid sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  Hid ds(H5Dcreate2(g, dname, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok() || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(3, std::string("dataset ") + gname + "/" + dname);
  if (write_scalar_attr(ds, "Time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &time) < 0)
    return fail(3, "attribute Time");
  return 0;
}
}  // namespace

int dgpu_h5_write_grid(const char* fname, const double* gy, int32_t n) {
  quiet();
  if (!fname || !gy || n < 2) return fail(1, "bad argument");
  Hid f(H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(3, std::string("cannot create ") + fname);
  Hid g(H5Gcreate2(f, "grids", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail(3, "group /grids");
  const hsize_t dims[1] = {(hsize_t)n};
  Hid sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
  Hid ds(H5Dcreate2(g, "y", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok() || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, gy) < 0)
    return fail(3, "dataset /grids/y");
  return 0;
}

int dgpu_h5_read_grid(const char* fname, double* gy, int32_t* n) {
  quiet();
  if (!fname || !n) return fail(1, "null argument");
  Hid f(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(3, std::string("cannot open ") + fname);
  Hid ds(H5Dopen2(f, "/grids/y", H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(3, "no dataset /grids/y");
  Hid sp(H5Dget_space(ds), H5Sclose);
  hsize_t dims[1] = {0};
  if (!sp.ok() || H5Sget_simple_extent_ndims(sp) != 1 || H5Sget_simple_extent_dims(sp, dims, nullptr) != 1)
    return fail(3, "/grids/y is not one-dimensional");
  *n = (int32_t)dims[0];
  if (gy && H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, gy) < 0) return fail(3, "read /grids/y");
  return 0;
}

int dgpu_h5_append_profile(const char* fname, const char* gname, const double* diag, int32_t ny, double time) {
  if (!fname || !gname || !diag || ny <= 0) return fail(1, "dgpu_h5_append_profile: bad argument");
  const hsize_t n = (hsize_t)ny;
  return append_series(fname, gname, diag, 1, &n, time);
}

int dgpu_h5_append_plane(const char* fname, const char* gname, const double* var2d, int32_t n1, int32_t n2,
                         double time) {
  if (!fname || !gname || !var2d || n1 <= 0 || n2 <= 0) return fail(1, "dgpu_h5_append_plane: bad argument");
  const hsize_t dims[2] = {(hsize_t)n2, (hsize_t)n1};  // Fortran (n1, n2)
  return append_series(fname, gname, var2d, 2, dims, time);
}

int dgpu_h5_append_real(const char* fname, const char* gname, double value) {
  if (!fname || !gname) return fail(1, "dgpu_h5_append_real: bad argument");
  quiet();
  Hid f(open_series_file(fname), H5Fclose);
  if (!f.ok()) return fail(3, std::string("cannot open ") + fname);
  hid_t gid = -1;
  const int nsamp = series_sample(f, gname, gid);
  if (nsamp < 0) return fail(3, std::string("series group ") + gname);
  Hid g(gid, H5Gclose);
  char dname[32];
  std::snprintf(dname, sizeof dname, "%04d", nsamp);
  const hsize_t one = 1;
  Hid sp(H5Screate_simple(1, &one, nullptr), H5Sclose);
  Hid ds(H5Dcreate2(g, dname, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok() || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &value) < 0)
    return fail(3, std::string("dataset ") + gname + "/" + dname);
  return 0;
}

int dgpu_write_flow(dgpu_handle* h, const char* fname, const dgpu_flow_times* t, int32_t save_pressure) {
  int32_t nx, ny, nz, nth;
  if (!h || !fname || !t) return fail(1, "dgpu_write_flow: null argument");
  if (dgpu_get_sizes(h, &nx, &ny, &nz, &nth)) return fail(5, dgpu_last_error());
  const Dims d{nx, ny, nz};
  const size_t n = d.field();
  std::vector<double> cu1(n), cu2(n), cu3(n), cp(n), cth(n * (nth > 0 ? nth : 1));
  std::vector<double> u1(n), u2(n), u3(n), th(n * (nth > 0 ? nth : 1)), p;
  auto dev = [&](int rc) { return rc ? fail(5, dgpu_last_error()) : 0; };
  int rc;
  if ((rc = dev(dgpu_download_state(h, cu1.data(), cu2.data(), cu3.data(), cp.data(), cth.data())))) return rc;
  // fft_xz_to_physical of the state (for/phdf5.f90:1083-1088, 1147)
  if ((rc = dev(dgpu_fft_xz_to_physical(h, cu1.data(), u1.data()))) ||
      (rc = dev(dgpu_fft_xz_to_physical(h, cu2.data(), u2.data()))) ||
      (rc = dev(dgpu_fft_xz_to_physical(h, cu3.data(), u3.data()))))
    return rc;
  for (int m = 0; m < nth; ++m)
    if ((rc = dev(dgpu_fft_xz_to_physical(h, cth.data() + m * n, th.data() + m * n)))) return rc;
  if (save_pressure) {
    p.resize(n);
    if ((rc = dev(dgpu_fft_xz_to_physical(h, cp.data(), p.data())))) return rc;
  }
  dgpu_g2gf(u2.data(), nx, ny, nz);
  const int wrc = dgpu_h5_write_flow(fname, nx, ny, nz, nth, u1.data(), u2.data(), u3.data(), th.data(),
                                     save_pressure ? p.data() : nullptr, t, nullptr);
  const std::string werr = g_err;
  dgpu_gf2g(u2.data(), nx, ny, nz);
  // the closing transforms back to Fourier space (for/phdf5.f90:1171, 1187-1192)
  if ((rc = dev(dgpu_fft_xz_to_fourier(h, u1.data(), cu1.data()))) ||
      (rc = dev(dgpu_fft_xz_to_fourier(h, u2.data(), cu2.data()))) ||
      (rc = dev(dgpu_fft_xz_to_fourier(h, u3.data(), cu3.data()))))
    return rc;
  for (int m = 0; m < nth; ++m)
    if ((rc = dev(dgpu_fft_xz_to_fourier(h, th.data() + m * n, cth.data() + m * n)))) return rc;
  if (save_pressure && (rc = dev(dgpu_fft_xz_to_fourier(h, p.data(), cp.data())))) return rc;
  if ((rc = dev(dgpu_upload_state(h, cu1.data(), cu2.data(), cu3.data(), cp.data(), cth.data())))) return rc;
  if (wrc) return fail(wrc, werr);
  return 0;
}

int dgpu_read_flow(dgpu_handle* h, const char* fname, uint32_t th_mask, dgpu_flow_times* t,
                   int32_t* read_pressure) {
  int32_t nx, ny, nz, nth;
  if (!h || !fname || !t || !read_pressure) return fail(1, "dgpu_read_flow: null argument");
  if (dgpu_get_sizes(h, &nx, &ny, &nz, &nth)) return fail(5, dgpu_last_error());
  const Dims d{nx, ny, nz};
  const size_t n = d.field();
  std::vector<double> cu1(n), cu2(n), cu3(n), cp(n), cth(n * (nth > 0 ? nth : 1));
  std::vector<double> u1(n), u2(n), u3(n), th(n * (nth > 0 ? nth : 1), 0.0), p(n);
  auto dev = [&](int rc) { return rc ? fail(5, dgpu_last_error()) : 0; };
  int rc;
  // the state the file does not replace (scalars outside th_mask, cp without P)
  if ((rc = dev(dgpu_download_state(h, cu1.data(), cu2.data(), cu3.data(), cp.data(), cth.data())))) return rc;
  int32_t has_p = 0;
  if ((rc = dgpu_h5_read_flow(fname, nx, ny, nz, nth, th_mask, u1.data(), u2.data(), u3.data(), th.data(),
                              p.data(), &has_p, t)))
    return rc;
  dgpu_gf2g(u2.data(), nx, ny, nz);
  // with variable_dt, courant on the velocity read (for/phdf5.f90:1437-1439):
  // the time loop's next dgpu_courant returns its delta_t
  if ((rc = dev(dgpu_courant_physical(h, u1.data(), u2.data(), u3.data())))) return rc;
  if ((rc = dev(dgpu_fft_xz_to_fourier(h, u1.data(), cu1.data()))) ||
      (rc = dev(dgpu_fft_xz_to_fourier(h, u2.data(), cu2.data()))) ||
      (rc = dev(dgpu_fft_xz_to_fourier(h, u3.data(), cu3.data()))))
    return rc;
  for (int m = 0; m < nth; ++m)
    if (((th_mask >> m) & 1u) && (rc = dev(dgpu_fft_xz_to_fourier(h, th.data() + m * n, cth.data() + m * n))))
      return rc;
  if (has_p && (rc = dev(dgpu_fft_xz_to_fourier(h, p.data(), cp.data())))) return rc;
  if ((rc = dev(dgpu_upload_state(h, cu1.data(), cu2.data(), cu3.data(), cp.data(), cth.data())))) return rc;
  // without P the reference recomputes the pressure (for/phdf5.f90:1452-1454)
  if (!has_p && (rc = dev(dgpu_poisson_p(h)))) return rc;
  *read_pressure = has_p;
  return 0;
}

}  // extern "C"
