// HDF5 / XDMF output of the HIP backend's solver state in libmpdata++'s
// on-disk layout (output/hdf5.hpp, output/hdf5_xdmf.hpp,
// output/detail/xdmf_writer.hpp).  Host code over the HDF5 C library; the
// C ABI is declared in include/mpdx_h5.h.
//
// What has to match the reference, and why it is written this way:
//  - dataset and attribute names, types, shapes, chunking and the deflate
//    level, so h5diff of two runs compares data, and readers written for the
//    reference's files (Paraview's XDMF reader, the reference's Python
//    plotting scripts) open these;
//  - the float conversions: coordinates and times are computed in the
//    solver's real_t and converted to float32 by HDF5, as the reference
//    writes them (hdf5.hpp:166-215), so X/Y/Z/T compare bit for bit;
//  - the XDMF text: element order, attribute order, tab indentation and the
//    "%f" time stamp of boost::property_tree's write_xml as the reference
//    calls it (xdmf_writer.hpp:156-204), so the .xmf files compare byte for
//    byte.
#include <hdf5.h>

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpdx_h5.h"

// built with -fvisibility=hidden: only the C ABI is exported
#define MPDX_H5_API __attribute__((visibility("default")))

namespace {

thread_local std::string g_err;

int fail(const std::string &msg)
{
  g_err = msg;
  return 1;
}

// Suppresses HDF5's automatic error-stack printing for the duration of one
// call; failures are reported through return codes and mpdx_h5_last_error.
struct QuietH5
{
  H5E_auto2_t fn = nullptr;
  void *data = nullptr;
  QuietH5()
  {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// RAII holder of an HDF5 identifier with its matching close routine.
struct Hid
{
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  ~Hid()
  {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
  operator hid_t() const { return id; }
};

hid_t mem_type(int real_bytes) { return real_bytes == 8 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT; }

// a value of the solver's real_t, as the reference holds it before writing
double as_real(double v, int real_bytes) { return real_bytes == 8 ? v : double(float(v)); }

std::string base_name(int64_t timestep)
{
  // hdf5.hpp:237-242: "timestep" + setw(10) setfill('0')
  char buf[64];
  std::snprintf(buf, sizeof buf, "timestep%010lld", (long long)timestep);
  return buf;
}

int check_setup(const mpdx_h5_setup *s)
{
  if (!s) return fail("mpdx_h5: null setup");
  if (s->n_dims != 2 && s->n_dims != 3) return fail("mpdx_h5: only 2D and 3D output supported");
  if (s->real_bytes != 4 && s->real_bytes != 8) return fail("mpdx_h5: real_bytes must be 4 or 8");
  for (int d = 0; d < s->n_dims; ++d)
    if (s->grid[d] < 1) return fail("mpdx_h5: grid sizes must be positive");
  if (s->outwindow < 1) return fail("mpdx_h5: outwindow must be >= 1");
  if (!(s->outfreq > 0)) return fail("mpdx_h5: outfreq must be positive");
  return 0;
}

// opts_string (opts.hpp:66-78): names in the reference's fixed order, '|'-joined
std::string opts_string(uint32_t o)
{
  struct Nm
  {
    uint32_t bit;
    const char *name;
  };
  // bits as in include/mpdx.h; div_2nd/div_3rd/fot have no bit in this backend
  const Nm order[] = {{1u << 6, "nug"}, {1u << 5, "iga"}, {1u << 1, "abs"}, {1u << 7, "dfl"},
                      {1u << 2, "tot"}, {1u << 0, "fct"}, {1u << 3, "pfc"}, {1u << 4, "npa"},
                      {1u << 8, "khn"}};
  std::string r;
  for (const Nm &n : order)
    if (o & n.bit) r += std::string("|") + n.name;
  return r.empty() ? "0" : r.substr(1);
}

int64_t n_out(const mpdx_h5_setup *s, double nt)
{
  // hdf5.hpp:199-203; integer arithmetic for a step count, real for var_dt
  if (s->var_dt)
  {
    const double v = ((nt - s->outstart) / s->outfreq + 1) * s->outwindow + (s->outstart == 0 ? 0 : 1) -
                     std::max(0, s->outwindow - int(std::fmod(nt, s->outfreq)) - 1);
    return v < 0 ? 0 : int64_t(v);
  }
  const int64_t nti = int64_t(nt), of = int64_t(s->outfreq), os = int64_t(s->outstart);
  const int64_t v = ((nti - os) / of + 1) * s->outwindow + (os == 0 ? 0 : 1) -
                    std::max<int64_t>(0, s->outwindow - nti % of - 1);
  return v < 0 ? 0 : v;
}

int write_attr(hid_t loc, const char *name, hid_t ftype, hid_t mtype, const void *buf)
{
  const hsize_t one = 1;
  Hid sp(H5Screate_simple(1, &one, nullptr), H5Sclose);
  Hid at(H5Acreate2(loc, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!sp.ok() || !at.ok() || H5Awrite(at, mtype, buf) < 0)
    return fail(std::string("mpdx_h5: cannot write attribute ") + name);
  return 0;
}

int write_real_attr(hid_t loc, const char *name, double v, int real_bytes)
{
  const float vf = float(v);
  return write_attr(loc, name, H5T_NATIVE_FLOAT, mem_type(real_bytes), real_bytes == 8 ? (const void *)&v
                                                                                       : (const void *)&vf);
}

// record_params for the mpdata family (hdf5.hpp:624-656)
int record_params(hid_t file, const mpdx_h5_setup *s)
{
  Hid g(H5Gcreate2(file, "advection", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail("mpdx_h5: cannot create group advection");
  const std::string os = opts_string(s->opts);
  {
    Hid st(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(st, os.size());
    if (write_attr(g, "opts", st, st, os.data())) return 1;
  }
  // real-valued attributes: float32 in the file, converted by HDF5 from the
  // solver's real_t (hdf5.hpp:635-640)
  if (write_real_attr(g, "dt", s->dt, s->real_bytes)) return 1;
  const char *names[3] = {"di", "dj", "dk"};
  for (int d = 0; d < s->n_dims; ++d)
    if (write_real_attr(g, names[d], s->dijk[d], s->real_bytes)) return 1;
  {
    const hbool_t v = s->var_dt ? 1 : 0;
    if (write_attr(g, "var_dt", H5T_NATIVE_HBOOL, H5T_NATIVE_HBOOL, &v)) return 1;
  }
  if (s->var_dt && write_real_attr(g, "max_courant", s->max_courant, s->real_bytes)) return 1;
  {
    const int v = s->n_iters;
    if (write_attr(g, "n_iters", H5T_NATIVE_INT, H5T_NATIVE_INT, &v)) return 1;
  }
  return 0;
}

// a contiguous float32 dataset filled from a real_t buffer
int write_plain(hid_t file, const char *name, int rank, const hsize_t *shape, const void *buf, int real_bytes)
{
  Hid sp(H5Screate_simple(rank, shape, nullptr), H5Sclose);
  Hid ds(H5Dcreate2(file, name, H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!sp.ok() || !ds.ok() || H5Dwrite(ds, mem_type(real_bytes), H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
    return fail(std::string("mpdx_h5: cannot write dataset ") + name);
  return 0;
}

// coordinates d * index over the (n+1)^n_dims corner points (hdf5.hpp:166-197)
template <typename real_t>
std::vector<real_t> coords(const mpdx_h5_setup *s, int dim)
{
  const int n0 = s->grid[0] + 1, n1 = s->grid[1] + 1, n2 = s->n_dims == 3 ? s->grid[2] + 1 : 1;
  std::vector<real_t> c(size_t(n0) * n1 * n2);
  const real_t d = real_t(s->dijk[dim]);
  size_t q = 0;
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k) c[q++] = d * real_t(dim == 0 ? i : dim == 1 ? j : k);
  return c;
}

// output times (hdf5.hpp:205-215), in the solver's real_t
template <typename real_t>
std::vector<real_t> times(const mpdx_h5_setup *s, int64_t n)
{
  std::vector<real_t> t(size_t(n > 0 ? n : 0));
  const real_t scale = s->var_dt ? real_t(1) : real_t(s->dt);
  const int ow = s->outwindow;
  auto f = [&](int64_t i) {
    return real_t(double(scale) * (double(i / ow) * s->outfreq + double(i % ow)));
  };
  if (s->outstart == 0)
    for (int64_t i = 0; i < n; ++i) t[i] = f(i);
  else if (n > 0)
  {
    t[0] = 0;
    for (int64_t j = 0; j + 1 < n; ++j)
      t[j + 1] = real_t(double(scale) * (double(j / ow) * s->outfreq + double(j % ow) + s->outstart));
  }
  return t;
}

// ---- XDMF text (xdmf_writer.hpp) ----

std::string xml_escape(const std::string &v)
{
  std::string r;
  for (char c : v)
  {
    switch (c)
    {
      case '&': r += "&amp;"; break;
      case '<': r += "&lt;"; break;
      case '>': r += "&gt;"; break;
      case '"': r += "&quot;"; break;
      case '\'': r += "&apos;"; break;
      default: r += c;
    }
  }
  return r;
}

std::string dims_string(const mpdx_h5_setup *s, int plus)
{
  std::string r;
  for (int d = 0; d < s->n_dims; ++d) r += std::to_string(s->grid[d] + plus) + " ";
  return r;
}

int write_text(const std::string &path, const std::string &text)
{
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return fail("mpdx_h5: cannot open " + path + ": " + std::strerror(errno));
  const size_t w = std::fwrite(text.data(), 1, text.size(), f);
  const int rc = std::fclose(f);
  if (w != text.size() || rc != 0) return fail("mpdx_h5: cannot write " + path);
  return 0;
}

std::string xmf_grid(const mpdx_h5_setup *s, const std::string &h5name, double time,
                     std::vector<std::string> names)
{
  // one Uniform grid: Time, Topology, Geometry over const.h5, then the
  // attributes sorted by name (they live in a std::set in the reference)
  std::sort(names.begin(), names.end());
  names.erase(std::unique(names.begin(), names.end()), names.end());
  char tbuf[64];
  std::snprintf(tbuf, sizeof tbuf, "%f", time);  // std::to_string(double)
  const std::string nd1 = dims_string(s, 1), nd0 = dims_string(s, 0);
  const char *xyz[3] = {"X", "Y", "Z"};
  std::string x = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<Xdmf>\n\t<Domain>\n";
  x += "\t\t<Grid Name=\"Grid\" GridType=\"Uniform\" xml:id=\"gid\">\n";
  x += std::string("\t\t\t<Time Value=\"") + tbuf + "\"/>\n";
  x += std::string("\t\t\t<Topology TopologyType=\"") + (s->n_dims == 2 ? "2DSMesh" : "3DSMesh") +
       "\" Dimensions=\"" + nd1 + "\"/>\n";
  x += std::string("\t\t\t<Geometry GeometryType=\"") + (s->n_dims == 2 ? "X_Y" : "X_Y_Z") + "\">\n";
  for (int d = 0; d < s->n_dims; ++d)
    x += "\t\t\t\t<DataItem Dimensions=\"" + nd1 + "\" NumberType=\"Float\" Format=\"HDF\">const.h5:/" + xyz[d] +
         "</DataItem>\n";
  x += "\t\t\t</Geometry>\n";
  for (const std::string &n : names)
  {
    x += "\t\t\t<Attribute Name=\"" + xml_escape(n) + "\" AttributeType=\"Scalar\" Center=\"Cell\">\n";
    x += "\t\t\t\t<DataItem Dimensions=\"" + nd0 + "\" NumberType=\"Float\" Format=\"HDF\">" +
         xml_escape(h5name + ":/" + n) + "</DataItem>\n";
    x += "\t\t\t</Attribute>\n";
  }
  x += "\t\t</Grid>\n\t</Domain>\n</Xdmf>\n";
  return x;
}

std::string xmf_temporal(const std::vector<int64_t> &steps)
{
  std::string x = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n";
  x += "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\">\n\t<Domain>\n";
  x += "\t\t<Grid Name=\"TimeGrid\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
  for (int64_t t : steps) x += "\t\t\t<xi:include href=\"" + base_name(t) + ".xmf\" xpointer=\"gid\"/>\n";
  x += "\t\t</Grid>\n\t</Domain>\n</Xdmf>\n";
  return x;
}

}  // namespace

extern "C" {

MPDX_H5_API const char *mpdx_h5_last_error(void) { return g_err.c_str(); }

MPDX_H5_API int64_t mpdx_h5_n_out(const mpdx_h5_setup *s, double nt)
{
  if (check_setup(s)) return -1;
  return n_out(s, nt);
}

MPDX_H5_API int mpdx_h5_start(const mpdx_h5_setup *s, const char *outdir, double nt)
{
  if (check_setup(s)) return 1;
  if (!outdir) return fail("mpdx_h5: null outdir");
  if (s->var_dt && s->outstart != 0) return fail("mpdx_h5: outstart must be 0 with var_dt");
  QuietH5 quiet;
  // boost::filesystem::create_directory: one level, an existing one is fine
  if (mkdir(outdir, 0777) != 0 && errno != EEXIST)
    return fail(std::string("mpdx_h5: cannot create ") + outdir + ": " + std::strerror(errno));
  const std::string path = std::string(outdir) + "/const.h5";
  Hid f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail("mpdx_h5: cannot create " + path);
  hsize_t cshape[3];
  for (int d = 0; d < s->n_dims; ++d) cshape[d] = hsize_t(s->grid[d]) + 1;
  const char *xyz[3] = {"X", "Y", "Z"};
  for (int d = 0; d < s->n_dims; ++d)
  {
    int rc;
    if (s->real_bytes == 8)
    {
      const auto c = coords<double>(s, d);
      rc = write_plain(f, xyz[d], s->n_dims, cshape, c.data(), 8);
    }
    else
    {
      const auto c = coords<float>(s, d);
      rc = write_plain(f, xyz[d], s->n_dims, cshape, c.data(), 4);
    }
    if (rc) return rc;
  }
  const int64_t n = n_out(s, nt);
  const hsize_t tn = hsize_t(n);
  if (s->real_bytes == 8)
  {
    const auto t = times<double>(s, n);
    if (write_plain(f, "T", 1, &tn, t.data(), 8)) return 1;
  }
  else
  {
    const auto t = times<float>(s, n);
    if (write_plain(f, "T", 1, &tn, t.data(), 4)) return 1;
  }
  if (record_params(f, s)) return 1;
  return 0;
}

MPDX_H5_API int mpdx_h5_record(const mpdx_h5_setup *s, const char *outdir, int64_t timestep, double record_time, int n_vars,
                   const char *const *names, const void *const *fields, int n_prev,
                   const int64_t *prev_timesteps)
{
  if (check_setup(s)) return 1;
  if (!outdir) return fail("mpdx_h5: null outdir");
  if (n_vars < 0 || (n_vars > 0 && (!names || !fields))) return fail("mpdx_h5: bad variable list");
  if (n_prev < 0 || (n_prev > 0 && !prev_timesteps)) return fail("mpdx_h5: bad list of earlier records");
  for (int v = 0; v < n_vars; ++v)
    if (!names[v] || !fields[v]) return fail("mpdx_h5: null variable name or field");
  QuietH5 quiet;
  const std::string dir(outdir), base = base_name(timestep);
  // hdf5_xdmf.hpp:64-80: the grid markup, then the temporal collection
  if (s->xdmf)
  {
    std::vector<std::string> nv(names, names + n_vars);
    if (write_text(dir + "/" + base + ".xmf", xmf_grid(s, base + ".h5", as_real(record_time, s->real_bytes), nv)))
      return 1;
    std::vector<int64_t> steps(prev_timesteps, prev_timesteps + n_prev);
    steps.push_back(timestep);
    if (write_text(dir + "/temp.xmf", xmf_temporal(steps))) return 1;
  }

  // hdf5.hpp:256-282: float32 datasets, one chunk of the whole grid, deflate 5
  const std::string path = dir + "/" + base + ".h5";
  Hid f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail("mpdx_h5: cannot create " + path);
  hsize_t shape[3], chunk[3];
  for (int d = 0; d < s->n_dims; ++d) shape[d] = chunk[d] = hsize_t(s->grid[d]);
  // HDF5 limits a chunk to 4 GiB; beyond that split the slowest dimension
  while (chunk[0] > 1 && double(chunk[0]) * chunk[1] * (s->n_dims == 3 ? chunk[2] : 1) * 4.0 >= 4294967296.0)
    chunk[0] = (chunk[0] + 1) / 2;
  Hid dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
  if (!dcpl.ok() || H5Pset_chunk(dcpl, s->n_dims, chunk) < 0 || H5Pset_deflate(dcpl, 5) < 0)
    return fail("mpdx_h5: cannot set chunking / deflate");
  Hid sp(H5Screate_simple(s->n_dims, shape, nullptr), H5Sclose);
  for (int v = 0; v < n_vars; ++v)
  {
    Hid ds(H5Dcreate2(f, names[v], H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose);
    if (!ds.ok() || H5Dwrite(ds, mem_type(s->real_bytes), H5S_ALL, H5S_ALL, H5P_DEFAULT, fields[v]) < 0)
      return fail(std::string("mpdx_h5: cannot write dataset ") + names[v] + " to " + path);
  }
  return 0;
}

MPDX_H5_API int mpdx_h5_read(const char *path, const char *name, float *out, int64_t n)
{
  if (!path || !name || (!out && n > 0)) return fail("mpdx_h5_read: null argument");
  QuietH5 quiet;
  Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(std::string("mpdx_h5_read: cannot open ") + path);
  Hid ds(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(std::string("mpdx_h5_read: no dataset ") + name + " in " + path);
  Hid sp(H5Dget_space(ds), H5Sclose);
  const hssize_t np = sp.ok() ? H5Sget_simple_extent_npoints(sp) : -1;
  if (np != n)
    return fail(std::string("mpdx_h5_read: dataset ") + name + " holds " + std::to_string((long long)np) +
                " values, " + std::to_string((long long)n) + " requested");
  if (H5Dread(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    return fail(std::string("mpdx_h5_read: cannot read ") + name);
  return 0;
}

}  // extern "C"
