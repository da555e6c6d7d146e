// HDF5 primitives for the instance files, over libhdf5 (1.10 C API) loaded with dlopen.
// Only the C prototypes come from hdf5.h (through decltype); the library is resolved at
// run time, so nothing here is linked against it.
#include "qed_h5.hpp"

#include <dlfcn.h>
#include <sys/stat.h>
#include <hdf5.h>

#include <cmath>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <type_traits>

namespace qed {
namespace h5 {
namespace {

struct Api {
  decltype(&H5open) open;
  decltype(&H5Eset_auto2) Eset_auto2;
  decltype(&H5Fcreate) Fcreate;
  decltype(&H5Fopen) Fopen;
  decltype(&H5Fclose) Fclose;
  decltype(&H5Gcreate2) Gcreate2;
  decltype(&H5Gopen2) Gopen2;
  decltype(&H5Gclose) Gclose;
  decltype(&H5Lexists) Lexists;
  decltype(&H5Ldelete) Ldelete;
  decltype(&H5Screate) Screate;
  decltype(&H5Screate_simple) Screate_simple;
  decltype(&H5Sget_simple_extent_ndims) Sget_simple_extent_ndims;
  decltype(&H5Sget_simple_extent_dims) Sget_simple_extent_dims;
  decltype(&H5Sclose) Sclose;
  decltype(&H5Tcopy) Tcopy;
  decltype(&H5Tset_size) Tset_size;
  decltype(&H5Tget_size) Tget_size;
  decltype(&H5Tis_variable_str) Tis_variable_str;
  decltype(&H5Tclose) Tclose;
  decltype(&H5Acreate2) Acreate2;
  decltype(&H5Aopen) Aopen;
  decltype(&H5Aexists) Aexists;
  decltype(&H5Awrite) Awrite;
  decltype(&H5Aread) Aread;
  decltype(&H5Aget_type) Aget_type;
  decltype(&H5Aclose) Aclose;
  decltype(&H5Dcreate2) Dcreate2;
  decltype(&H5Dopen2) Dopen2;
  decltype(&H5Dwrite) Dwrite;
  decltype(&H5Dread) Dread;
  decltype(&H5Dget_space) Dget_space;
  decltype(&H5Dclose) Dclose;
  decltype(&H5free_memory) free_memory;
  hid_t t_double = -1, t_int = -1, t_c_s1 = -1;
};

Api* load() {
  static std::once_flag once;
  static Api* api = nullptr;
  static std::string why;
  std::call_once(once, [] {
    const char* names[] = {std::getenv("QED_HDF5_LIB"), "libhdf5.so.103", "libhdf5.so",
                           "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    void* h = nullptr;
    for (const char* n : names)
      if (n && *n && (h = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) {
      why = "the HDF5 library (libhdf5.so) was not found; set QED_HDF5_LIB to its path";
      return;
    }
    static Api a;
    bool ok = true;
    auto sym = [&](auto& fp, const char* s) {
      fp = reinterpret_cast<std::remove_reference_t<decltype(fp)>>(dlsym(h, s));
      if (!fp) {
        ok = false;
        why = std::string("symbol ") + s + " missing from the HDF5 library";
      }
    };
    sym(a.open, "H5open");
    sym(a.Eset_auto2, "H5Eset_auto2");
    sym(a.Fcreate, "H5Fcreate");
    sym(a.Fopen, "H5Fopen");
    sym(a.Fclose, "H5Fclose");
    sym(a.Gcreate2, "H5Gcreate2");
    sym(a.Gopen2, "H5Gopen2");
    sym(a.Gclose, "H5Gclose");
    sym(a.Lexists, "H5Lexists");
    sym(a.Ldelete, "H5Ldelete");
    sym(a.Screate, "H5Screate");
    sym(a.Screate_simple, "H5Screate_simple");
    sym(a.Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
    sym(a.Sget_simple_extent_dims, "H5Sget_simple_extent_dims");
    sym(a.Sclose, "H5Sclose");
    sym(a.Tcopy, "H5Tcopy");
    sym(a.Tset_size, "H5Tset_size");
    sym(a.Tget_size, "H5Tget_size");
    sym(a.Tis_variable_str, "H5Tis_variable_str");
    sym(a.Tclose, "H5Tclose");
    sym(a.Acreate2, "H5Acreate2");
    sym(a.Aopen, "H5Aopen");
    sym(a.Aexists, "H5Aexists");
    sym(a.Awrite, "H5Awrite");
    sym(a.Aread, "H5Aread");
    sym(a.Aget_type, "H5Aget_type");
    sym(a.Aclose, "H5Aclose");
    sym(a.Dcreate2, "H5Dcreate2");
    sym(a.Dopen2, "H5Dopen2");
    sym(a.Dwrite, "H5Dwrite");
    sym(a.Dread, "H5Dread");
    sym(a.Dget_space, "H5Dget_space");
    sym(a.Dclose, "H5Dclose");
    sym(a.free_memory, "H5free_memory");
    if (!ok) return;
    if (a.open() < 0) {
      why = "H5open failed";
      return;
    }
    // predefined type identifiers are library globals, valid after H5open
    auto* d = static_cast<hid_t*>(dlsym(h, "H5T_NATIVE_DOUBLE_g"));
    auto* i = static_cast<hid_t*>(dlsym(h, "H5T_NATIVE_INT_g"));
    auto* s = static_cast<hid_t*>(dlsym(h, "H5T_C_S1_g"));
    if (!d || !i || !s) {
      why = "predefined types missing from the HDF5 library";
      return;
    }
    a.t_double = *d;
    a.t_int = *i;
    a.t_c_s1 = *s;
    a.Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as status codes
    api = &a;
  });
  if (!api) throw std::runtime_error(why);
  return api;
}

[[noreturn]] void bad(const std::string& what, const std::string& name) {
  throw std::runtime_error(what + " failed for '" + name + "'");
}

// closes an HDF5 id of one kind when it leaves scope
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0) close(id);
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
};

hid_t vlen_string_type(Api* a) {
  hid_t t = a->Tcopy(a->t_c_s1);
  if (t < 0 || a->Tset_size(t, H5T_VARIABLE) < 0) throw std::runtime_error("H5Tcopy/H5Tset_size failed");
  return t;
}

}  // namespace

Node& Node::operator=(Node&& o) noexcept {
  if (this != &o) {
    this->~Node();
    id_ = o.id_;
    file_ = o.file_;
    o.id_ = -1;
  }
  return *this;
}

Node::~Node() {
  if (id_ < 0) return;
  Api* a = load();
  if (file_) a->Fclose(id_);
  else a->Gclose(id_);
  id_ = -1;
}

Node open_or_create(const std::string& path) {
  Api* a = load();
  struct stat st {};
  const bool there = stat(path.c_str(), &st) == 0;
  hid_t f = there ? a->Fopen(path.c_str(), 0x0001u /* read-write */, H5P_DEFAULT)
                  : a->Fcreate(path.c_str(), 0x0002u /* truncate */, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) bad(there ? "opening the file" : "creating the file", path);
  return Node(f, true);
}

Node open_read(const std::string& path) {
  Api* a = load();
  hid_t f = a->Fopen(path.c_str(), 0x0000u /* read-only */, H5P_DEFAULT);
  if (f < 0) bad("opening the file", path);
  return Node(f, true);
}

Node create_group(const Node& at, const std::string& name) {
  Api* a = load();
  hid_t g = a->Gcreate2(at.id(), name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) bad("creating group", name);
  return Node(g, false);
}

Node open_group(const Node& at, const std::string& name) {
  Api* a = load();
  hid_t g = a->Gopen2(at.id(), name.c_str(), H5P_DEFAULT);
  if (g < 0) bad("opening group", name);
  return Node(g, false);
}

bool exists(const Node& at, const std::string& name) {
  htri_t r = load()->Lexists(at.id(), name.c_str(), H5P_DEFAULT);
  if (r < 0) bad("looking up", name);
  return r > 0;
}

void unlink(const Node& at, const std::string& name) {
  if (load()->Ldelete(at.id(), name.c_str(), H5P_DEFAULT) < 0) bad("unlinking", name);
}

namespace {
void write_scalar_attr(const Node& at, const std::string& name, hid_t ftype, hid_t mtype, const void* v) {
  Api* a = load();
  Id sp(a->Screate(H5S_SCALAR), a->Sclose);
  if (sp.id < 0) bad("creating a scalar space for", name);
  Id at_(a->Acreate2(at.id(), name.c_str(), ftype, sp.id, H5P_DEFAULT, H5P_DEFAULT), a->Aclose);
  if (at_.id < 0) bad("creating attribute", name);
  if (a->Awrite(at_.id, mtype, v) < 0) bad("writing attribute", name);
}

void read_scalar_attr(const Node& at, const std::string& name, hid_t mtype, void* v) {
  Api* a = load();
  Id at_(a->Aopen(at.id(), name.c_str(), H5P_DEFAULT), a->Aclose);
  if (at_.id < 0) bad("opening attribute", name);
  if (a->Aread(at_.id, mtype, v) < 0) bad("reading attribute", name);
}
}  // namespace

void write_attr(const Node& at, const std::string& name, double v) {
  Api* a = load();
  write_scalar_attr(at, name, a->t_double, a->t_double, &v);
}

void write_attr(const Node& at, const std::string& name, int v) {
  Api* a = load();
  write_scalar_attr(at, name, a->t_int, a->t_int, &v);
}

void write_attr(const Node& at, const std::string& name, const std::string& v) {
  Api* a = load();
  Id t(vlen_string_type(a), a->Tclose);
  const char* p = v.c_str();
  write_scalar_attr(at, name, t.id, t.id, &p);
}

bool attr_exists(const Node& at, const std::string& name) {
  htri_t r = load()->Aexists(at.id(), name.c_str());
  if (r < 0) bad("looking up attribute", name);
  return r > 0;
}

double read_attr_double(const Node& at, const std::string& name) {
  double v = 0.0;
  read_scalar_attr(at, name, load()->t_double, &v);
  return v;
}

int read_attr_int(const Node& at, const std::string& name) {
  int v = 0;
  read_scalar_attr(at, name, load()->t_int, &v);
  return v;
}

std::string read_attr_string(const Node& at, const std::string& name) {
  Api* a = load();
  Id at_(a->Aopen(at.id(), name.c_str(), H5P_DEFAULT), a->Aclose);
  if (at_.id < 0) bad("opening attribute", name);
  Id ft(a->Aget_type(at_.id), a->Tclose);
  if (ft.id < 0) bad("reading the type of attribute", name);
  const htri_t vl = a->Tis_variable_str(ft.id);
  if (vl < 0) bad("reading the type of attribute", name);
  if (vl > 0) {
    Id mt(vlen_string_type(a), a->Tclose);
    char* p = nullptr;
    if (a->Aread(at_.id, mt.id, &p) < 0) bad("reading attribute", name);
    std::string s = p ? p : "";
    if (p) a->free_memory(p);
    return s;
  }
  // fixed-length string
  const size_t n = a->Tget_size(ft.id);
  Id mt(a->Tcopy(a->t_c_s1), a->Tclose);
  if (mt.id < 0 || a->Tset_size(mt.id, n + 1) < 0) bad("reading attribute", name);
  std::string s(n + 1, '\0');
  if (a->Aread(at_.id, mt.id, &s[0]) < 0) bad("reading attribute", name);
  s.resize(std::string(s.c_str()).size());
  return s;
}

void write_vec(const Node& at, const std::string& name, const std::vector<double>& v) {
  Api* a = load();
  hsize_t dims[1] = {(hsize_t)v.size()};
  Id sp(a->Screate_simple(1, dims, nullptr), a->Sclose);
  if (sp.id < 0) bad("creating a data space for", name);
  Id ds(a->Dcreate2(at.id(), name.c_str(), a->t_double, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), a->Dclose);
  if (ds.id < 0) bad("creating dataset", name);
  if (!v.empty() && a->Dwrite(ds.id, a->t_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    bad("writing dataset", name);
}

std::vector<double> read_vec(const Node& at, const std::string& name) {
  Api* a = load();
  Id ds(a->Dopen2(at.id(), name.c_str(), H5P_DEFAULT), a->Dclose);
  if (ds.id < 0) bad("opening dataset", name);
  Id sp(a->Dget_space(ds.id), a->Sclose);
  if (sp.id < 0 || a->Sget_simple_extent_ndims(sp.id) != 1) bad("reading the shape of dataset", name);
  hsize_t dims[1] = {0};
  a->Sget_simple_extent_dims(sp.id, dims, nullptr);
  std::vector<double> v((size_t)dims[0]);
  if (!v.empty() && a->Dread(ds.id, a->t_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    bad("reading dataset", name);
  return v;
}

void write_mat(const Node& at, const std::string& name, int r, int c, const double* v) {
  Node g = create_group(at, name);
  write_attr(g, "r", r);
  write_attr(g, "c", c);
  write_vec(g, "data", std::vector<double>(v, v + (size_t)r * c));
}

void write_mat_complex(const Node& at, const std::string& name, int r, int c, const double* re, const double* im) {
  Node g = create_group(at, name);
  write_attr(g, "r", r);
  write_attr(g, "c", c);
  const size_t n = (size_t)r * c;
  write_vec(g, "re", std::vector<double>(re, re + n));
  write_vec(g, "im", im ? std::vector<double>(im, im + n) : std::vector<double>(n, 0.0));
}

std::vector<double> read_mat_real(const Node& at, const std::string& name, int& r, int& c) {
  Node g = open_group(at, name);
  r = read_attr_int(g, "r");
  c = read_attr_int(g, "c");
  const size_t n = (size_t)r * c;
  std::vector<double> v;
  if (exists(g, "data")) {
    v = read_vec(g, "data");
  } else {
    v = read_vec(g, "re");
    const std::vector<double> im = read_vec(g, "im");
    for (double x : im)
      if (x != 0.0) throw std::runtime_error("matrix '" + name + "' is complex; this solver has a real Hilbert space");
  }
  if (v.size() != n) throw std::runtime_error("matrix '" + name + "' holds " + std::to_string(v.size()) +
                                              " values for " + std::to_string(r) + " x " + std::to_string(c));
  return v;
}

}  // namespace h5
}  // namespace qed
