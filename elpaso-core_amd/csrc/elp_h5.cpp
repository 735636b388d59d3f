// elPaSo HDF5 model input and result output (libelpaso_amd_h5.so).
//
// Reading follows cFemParserInterface::parseInputFile
// (misc/parser/femparserinterface.cpp:189-663) with the accessors of
// cFemParserHDF5 (misc/parser/femparserhdf5.cpp:62-560) and the names of
// misc/hdf5/rulesh5.h; the model is filled through the public libelpaso_amd
// calls.  Writing reproduces the datasets and attribute types of
// cOutputH5 (misc/hdf5/outputh5.cpp) as the frequency loop
// (analysis/frequency/analysisfrequencybasic.cpp:185-265) and
// MathLibrary::ExportPetscMatToHDF5Complex (math/mathlibrary.cpp:53-86) use it.
#include <hdf5.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/elpaso_amd_h5.h"

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
  int code;
  H5Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(const std::string &msg, int code = ELP_EINVAL) { throw H5Error(code, msg); }

// libelpaso_amd status -> exception carrying its message
void check(int rc, const std::string &what) {
  if (rc != ELP_OK) fail(what + ": " + elp_last_error(nullptr), rc);
}

// HDF5 handle closed on scope exit
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid() = default;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  Hid(Hid &&o) noexcept : id(o.id), close(o.close) { o.id = -1; }
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

Hid open_file(const char *path, unsigned flags) {
  if (!path) fail("null path");
  hid_t f = H5Fopen(path, flags, H5P_DEFAULT);
  if (f < 0) fail(std::string("cannot open HDF5 file ") + path);
  return Hid(f, H5Fclose);
}

bool exists(hid_t loc, const std::string &path) {
  // every component must exist for H5Lexists on the full path
  size_t pos = 0;
  while (true) {
    size_t next = path.find('/', pos + 1);
    std::string part = path.substr(0, next);
    if (!part.empty() && part != "/" && H5Lexists(loc, part.c_str(), H5P_DEFAULT) <= 0) return false;
    if (next == std::string::npos) return true;
    pos = next;
  }
}

Hid open_group(hid_t f, const std::string &name) {
  hid_t g = H5Gopen2(f, name.c_str(), H5P_DEFAULT);
  if (g < 0) fail("missing group " + name);
  return Hid(g, H5Gclose);
}

bool has_attr(hid_t obj, const char *name) { return H5Aexists(obj, name) > 0; }

// string attribute, variable or fixed length
std::string attr_string(hid_t obj, const char *name) {
  Hid a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (a < 0) fail(std::string("missing attribute ") + name);
  Hid ft(H5Aget_type(a), H5Tclose);
  if (H5Tget_class(ft) != H5T_STRING) fail(std::string("attribute ") + name + " is not a string");
  if (H5Tis_variable_str(ft) > 0) {
    Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, H5T_VARIABLE);
    H5Tset_cset(mt, H5Tget_cset(ft));  // no ASCII <-> UTF-8 conversion
    char *s = nullptr;
    if (H5Aread(a, mt, &s) < 0) fail(std::string("cannot read attribute ") + name);
    std::string r = s ? s : "";
    if (s) H5free_memory(s);
    return r;
  }
  size_t sz = H5Tget_size(ft);
  std::vector<char> buf(sz + 1, 0);
  Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(mt, sz + 1);
  H5Tset_cset(mt, H5Tget_cset(ft));
  if (H5Aread(a, mt, buf.data()) < 0) fail(std::string("cannot read attribute ") + name);
  return std::string(buf.data());
}

// numeric attribute; elPaSo stores some numbers as strings ("343.0")
double attr_number(hid_t obj, const char *name) {
  Hid a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (a < 0) fail(std::string("missing attribute ") + name);
  Hid ft(H5Aget_type(a), H5Tclose);
  if (H5Tget_class(ft) == H5T_STRING) {
    std::string s = attr_string(obj, name);
    errno = 0;
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || errno != 0) fail(std::string("attribute ") + name + " is not a number: '" + s + "'");
    return v;
  }
  if (H5Tget_class(ft) == H5T_INTEGER) {
    int64_t v = 0;
    if (H5Aread(a, H5T_NATIVE_INT64, &v) < 0) fail(std::string("cannot read attribute ") + name);
    return (double)v;
  }
  double v = 0;
  if (H5Aread(a, H5T_NATIVE_DOUBLE, &v) < 0) fail(std::string("cannot read attribute ") + name);
  return v;
}

int64_t attr_int(hid_t obj, const char *name) {
  Hid a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (a < 0) fail(std::string("missing attribute ") + name);
  Hid ft(H5Aget_type(a), H5Tclose);
  if (H5Tget_class(ft) == H5T_STRING) return (int64_t)attr_number(obj, name);
  int64_t v = 0;
  if (H5Aread(a, H5T_NATIVE_INT64, &v) < 0) fail(std::string("cannot read attribute ") + name);
  return v;
}

// names of a group's members in the index order elPaSo's getMemberName uses
std::vector<std::string> members(hid_t g) {
  H5G_info_t info;
  if (H5Gget_info(g, &info) < 0) fail("cannot list group");
  std::vector<std::string> out;
  for (hsize_t i = 0; i < info.nlinks; i++) {
    ssize_t len = H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
    if (len < 0) fail("cannot read a member name");
    std::string s(len + 1, '\0');
    H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, &s[0], len + 1, H5P_DEFAULT);
    s.resize(len);
    out.push_back(s);
  }
  return out;
}

// 2-D integer dataset -> rows x cols int64
std::vector<int64_t> read_int_matrix(hid_t d, int64_t &rows, int64_t &cols, const std::string &what) {
  Hid sp(H5Dget_space(d), H5Sclose);
  int nd = H5Sget_simple_extent_ndims(sp);
  if (nd != 2 && nd != 1) fail(what + ": expected a matrix");
  hsize_t dims[2] = {0, 1};
  H5Sget_simple_extent_dims(sp, dims, nullptr);
  rows = (int64_t)dims[0];
  cols = nd == 2 ? (int64_t)dims[1] : 1;
  std::vector<int64_t> v((size_t)(rows * cols));
  if (!v.empty() && H5Dread(d, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    fail(what + ": cannot read");
  return v;
}

void read_analysis(hid_t f, elp_h5_analysis *a) {
  Hid g = open_group(f, "/Analysis");
  std::memset(a, 0, sizeof(*a));
  a->revision = has_attr(g, "revision") ? attr_int(g, "revision") : 0;
  if (a->revision < 6) fail("input file revision " + std::to_string(a->revision) + " is older than 6");
  std::string t = has_attr(g, "type") ? attr_string(g, "type") : "";
  std::snprintf(a->type, sizeof(a->type), "%s", t.c_str());
  std::string d = has_attr(g, "description") ? attr_string(g, "description") : "";
  std::snprintf(a->description, sizeof(a->description), "%s", d.c_str());
  a->start = has_attr(g, "start") ? attr_number(g, "start") : 0.0;
  a->delta = has_attr(g, "delta") ? attr_number(g, "delta") : 0.0;
  a->steps = has_attr(g, "steps") ? attr_int(g, "steps") : 0;
  a->write_hdf5 = has_attr(g, "hdf5") ? (int32_t)attr_int(g, "hdf5") : 0;
}

// DOF enum order of the structure constraint flags (femparserhdf5.cpp:385-405)
const char *kConFlags[21] = {"u1", "u2",  "u3",  "w1",   "w2",    "w3",    "p0",    "p1",  "p2",  "p3",  "xd3",
                             "wd1", "wd2", "dwdx", "dwdy", "dwdxy", "fluid", "z_1", "z_3", "x1_2", "x2_2"};
constexpr int kFluidDof = 16;

int element_kind(const std::string &t, int &npe) {
  if (t == "Fluid8") return npe = 8, ELP_FLUID8;
  if (t == "Brick8") return npe = 8, ELP_BRICK8;
  if (t == "Fluid20") return npe = 20, ELP_FLUID20;
  if (t == "Brick20") return npe = 20, ELP_BRICK20;
  return 0;
}

void read_model(elp_model *m, hid_t f, elp_h5_counts *cnt) {
  // ---- nodes (femparserinterface.cpp:257-280)
  {
    if (!exists(f, "/Nodes/mtxFemNodes")) fail("missing dataset /Nodes/mtxFemNodes");
    Hid d(H5Dopen2(f, "/Nodes/mtxFemNodes", H5P_DEFAULT), H5Dclose);
    if (d < 0) fail("cannot open /Nodes/mtxFemNodes");
    Hid sp(H5Dget_space(d), H5Sclose);
    if (H5Sget_simple_extent_ndims(sp) != 1) fail("/Nodes/mtxFemNodes: expected a 1-D dataset of node records");
    hsize_t n = 0;
    H5Sget_simple_extent_dims(sp, &n, nullptr);
    struct Rec {
      int64_t id;
      double x, y, z;
    };
    Hid mt(H5Tcreate(H5T_COMPOUND, sizeof(Rec)), H5Tclose);
    H5Tinsert(mt, "Ids", HOFFSET(Rec, id), H5T_NATIVE_INT64);
    H5Tinsert(mt, "xCoords", HOFFSET(Rec, x), H5T_NATIVE_DOUBLE);
    H5Tinsert(mt, "yCoords", HOFFSET(Rec, y), H5T_NATIVE_DOUBLE);
    H5Tinsert(mt, "zCoords", HOFFSET(Rec, z), H5T_NATIVE_DOUBLE);
    std::vector<Rec> recs(n);
    if (n > 0 && H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, recs.data()) < 0)
      fail("cannot read /Nodes/mtxFemNodes (compound Ids, xCoords, yCoords, zCoords)");
    std::vector<int64_t> ids(n);
    std::vector<double> xyz(3 * n);
    for (hsize_t i = 0; i < n; i++) {
      ids[i] = recs[i].id;
      xyz[3 * i] = recs[i].x;
      xyz[3 * i + 1] = recs[i].y;
      xyz[3 * i + 2] = recs[i].z;
    }
    check(elp_model_set_nodes(m, (int64_t)n, ids.data(), xyz.data()), "/Nodes");
    cnt->nodes = (int64_t)n;
  }
  // ---- materials (femparserhdf5.cpp:213-326; materialfactory.cpp:41-100)
  std::map<int64_t, std::string> unsupported_mat;
  if (exists(f, "/Materials")) {
    Hid g = open_group(f, "/Materials");
    for (auto &name : members(g)) {
      Hid d(H5Oopen(g, name.c_str(), H5P_DEFAULT), H5Oclose);
      if (d < 0) fail("cannot open /Materials/" + name);
      const std::string type = attr_string(d, "MaterialType");
      const int64_t id = attr_int(d, "Id");
      const std::string where = "/Materials/" + name + " (" + type + ")";
      if (type == "AF_LIN_UAF_ISO_DIR") {
        check(elp_model_add_fluid_material(m, id, attr_number(d, "rho"), attr_number(d, "c")), where);
      } else if (type == "STR_LIN_ELA_ISO_DIR") {
        // elastic: no loss factor, so the visco type is irrelevant
        check(elp_model_add_structure_material(m, id, attr_number(d, "E"), attr_number(d, "nu"),
                                               attr_number(d, "rho"), 0.0, 1),
              where);
      } else if (type == "STR_LIN_VIS_ISO_DIR") {
        check(elp_model_add_structure_material(m, id, attr_number(d, "E"), attr_number(d, "nu"),
                                               attr_number(d, "rho"), attr_number(d, "eta"),
                                               (int)attr_int(d, "type")),
              where);
      } else {
        unsupported_mat[id] = type;  // an error only if an element group uses it
        continue;
      }
      cnt->materials++;
    }
  }
  // ---- element groups (femparserinterface.cpp:309-395, femparserhdf5.cpp:336-356)
  if (exists(f, "/Elements")) {
    Hid g = open_group(f, "/Elements");
    for (auto &name : members(g)) {
      Hid d(H5Dopen2(g, name.c_str(), H5P_DEFAULT), H5Dclose);
      if (d < 0) fail("cannot open /Elements/" + name);
      const std::string where = "/Elements/" + name;
      const std::string type = attr_string(d, "ElementType");
      int npe = 0;
      const int kind = element_kind(type, npe);
      if (!kind) fail(where + ": element type " + type + " is not supported (Fluid8, Brick8, Fluid20, Brick20)");
      const int64_t mat = attr_int(d, "MaterialId");
      if (mat <= 0) fail(where + ": no material defined for the group");
      if (unsupported_mat.count(mat)) fail(where + ": material type " + unsupported_mat[mat] + " is not supported");
      int64_t rows = 0, cols = 0;
      std::vector<int64_t> v = read_int_matrix(d, rows, cols, where);
      if (rows > 0 && cols != 1 + npe)
        fail(where + ": " + type + " rows need " + std::to_string(1 + npe) + " columns, found " + std::to_string(cols));
      if (has_attr(d, "N") && attr_int(d, "N") != rows)
        fail(where + ": attribute N differs from the number of rows");
      std::vector<int64_t> ids(rows), conn(rows * npe);
      for (int64_t e = 0; e < rows; e++) {
        ids[e] = v[e * cols];
        for (int k = 0; k < npe; k++) conn[e * npe + k] = v[e * cols + 1 + k];
      }
      check(elp_model_add_elements(m, kind, mat, rows, ids.data(), conn.data()), where);
      cnt->element_groups++;
      cnt->elements += rows;
    }
  }
  // ---- element loads (femparserinterface.cpp:397-423): normal velocity faces
  if (exists(f, "/ElemLoads")) {
    Hid g = open_group(f, "/ElemLoads");
    auto ls = members(g);
    if (!ls.empty()) fail("/ElemLoads/" + ls[0] + ": element loads are not supported");
  }
  // ---- node constraints (femparserinterface.cpp:535-565, boundaryconditionfactory.cpp:36-60)
  if (exists(f, "/NodeConstraints")) {
    Hid g = open_group(f, "/NodeConstraints");
    for (auto &name : members(g)) {
      Hid d(H5Oopen(g, name.c_str(), H5P_DEFAULT), H5Oclose);
      const std::string where = "/NodeConstraints/" + name;
      if (d < 0) fail("cannot open " + where);
      const std::string type = attr_string(d, "ConstraintType");
      const int64_t node = attr_int(d, "NodeId");
      if (type == "acoustic") {
        // the prescribed pressure is attribute "p" (sValP of rulesh5.h:332,
        // read by femparserhdf5.cpp:442), as elPaSo reads it
        check(elp_model_add_dirichlet(m, node, kFluidDof, attr_number(d, "p"), 0.0), where);
      } else if (type == "structure") {
        for (int k = 0; k < 21; k++) {
          if (!has_attr(d, kConFlags[k]) || attr_int(d, kConFlags[k]) == 0) continue;
          const std::string vname = std::string("val") + kConFlags[k];
          const double val = has_attr(d, vname.c_str()) ? attr_number(d, vname.c_str()) : 0.0;
          check(elp_model_add_dirichlet(m, node, k, val, 0.0), where);
        }
      } else {
        fail(where + ": constraint type " + type + " is not supported");
      }
      cnt->constraints++;
    }
  }
  // ---- node loads (femparserinterface.cpp:567-590, nodalforcefactory.cpp:36-55)
  if (exists(f, "/NodeLoads")) {
    Hid g = open_group(f, "/NodeLoads");
    for (auto &name : members(g)) {
      Hid d(H5Oopen(g, name.c_str(), H5P_DEFAULT), H5Oclose);
      const std::string where = "/NodeLoads/" + name;
      if (d < 0) fail("cannot open " + where);
      const std::string type = attr_string(d, "LoadType");
      const int64_t node = attr_int(d, "NodeId");
      if (type == "point_force") {
        check(elp_model_add_point_force(m, node, attr_number(d, "Fx"), attr_number(d, "Fy"), attr_number(d, "Fz")),
              where);
      } else if (type == "fluid") {
        check(elp_model_add_point_source(m, node, attr_number(d, "P")), where);
      } else {
        fail(where + ": load type " + type + " is not supported");
      }
      cnt->node_loads++;
    }
  }
  // ---- node moments are not part of a solid/fluid model
  if (exists(f, "/NodeMoments")) {
    Hid g = open_group(f, "/NodeMoments");
    auto ls = members(g);
    if (!ls.empty()) fail("/NodeMoments/" + ls[0] + ": node moments are not supported");
  }
  // ---- interface elements (femparserinterface.cpp:626-663, elementinterfacefactory.cpp:62-98)
  if (exists(f, "/InterfaceElements")) {
    Hid g = open_group(f, "/InterfaceElements");
    for (auto &name : members(g)) {
      Hid d(H5Dopen2(g, name.c_str(), H5P_DEFAULT), H5Dclose);
      const std::string where = "/InterfaceElements/" + name;
      if (d < 0) fail("cannot open " + where);
      const std::string type = attr_string(d, "type");
      if (type != "InterfaceMindlin") fail(where + ": interface type " + type + " is not supported");
      const int64_t nf = attr_int(d, "NfluidNodes"), ns = attr_int(d, "NstructNodes");
      if (nf != 4 || ns != 4) fail(where + ": only 4-node interface elements are supported");
      int64_t rows = 0, cols = 0;
      std::vector<int64_t> v = read_int_matrix(d, rows, cols, where);
      if (rows > 0 && cols != 2 + ns + nf) fail(where + ": rows need 10 columns");
      std::vector<int64_t> ids(rows), fn(rows * 4), sn(rows * 4);
      std::vector<double> ori(rows);
      for (int64_t e = 0; e < rows; e++) {
        const int64_t *r = &v[e * cols];
        ids[e] = r[0];
        ori[e] = (double)r[1];
        for (int k = 0; k < 4; k++) {
          sn[e * 4 + k] = r[2 + k];       // structure nodes first
          fn[e * 4 + k] = r[2 + ns + k];  // then fluid nodes
        }
      }
      check(elp_model_add_interfaces(m, rows, ids.data(), ori.data(), fn.data(), sn.data()), where);
      cnt->interfaces += rows;
    }
  }
  if (exists(f, "/NCInterfaceElements")) {
    Hid g = open_group(f, "/NCInterfaceElements");
    auto ls = members(g);
    if (!ls.empty()) fail("/NCInterfaceElements/" + ls[0] + ": non-conforming interfaces are not supported");
  }
}

// ---- writers (cOutputH5)

Hid create_or_open(const char *path, bool truncate) {
  if (!path) fail("null path");
  if (!truncate) {
    hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f >= 0) return Hid(f, H5Fclose);
  }
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) fail(std::string("cannot create HDF5 file ") + path);
  return Hid(f, H5Fclose);
}

Hid ensure_group(hid_t f, const std::string &name) {
  if (H5Lexists(f, name.c_str(), H5P_DEFAULT) > 0) return open_group(f, name);
  hid_t g = H5Gcreate2(f, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) fail("cannot create group " + name);
  return Hid(g, H5Gclose);
}

// compound {real, imag} of doubles (cOutputH5::appendComplexVector)
Hid complex_type() {
  Hid t(H5Tcreate(H5T_COMPOUND, 2 * sizeof(double)), H5Tclose);
  H5Tinsert(t, "real", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "imag", sizeof(double), H5T_NATIVE_DOUBLE);
  return t;
}

void write_dataset(hid_t loc, const std::string &name, hid_t type, int rank, const hsize_t *dims, const void *data) {
  Hid sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  Hid d(H5Dcreate2(loc, name.c_str(), type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (d < 0) fail("cannot create dataset " + name);
  if (H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) fail("cannot write dataset " + name);
}

// variable-length UTF-8 string attribute (cOutputH5::writeStringAttributeToGroup)
void write_string_attr(hid_t obj, const std::string &name, const std::string &value) {
  Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(t, H5T_VARIABLE);
  H5Tset_cset(t, H5T_CSET_UTF8);
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  Hid a(H5Acreate2(obj, name.c_str(), t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (a < 0) fail("cannot create attribute " + name);
  const char *s = value.c_str();
  if (H5Awrite(a, t, &s) < 0) fail("cannot write attribute " + name);
}

void write_int_attr(hid_t obj, const std::string &name, int value) {
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  Hid a(H5Acreate2(obj, name.c_str(), H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (a < 0) fail("cannot create attribute " + name);
  if (H5Awrite(a, H5T_NATIVE_INT, &value) < 0) fail("cannot write attribute " + name);
}

int32_t to_i32(int64_t v, const char *what) {
  if (v < INT32_MIN || v > INT32_MAX) fail(std::string(what) + " does not fit the int32 layout of the file");
  return (int32_t)v;
}

template <class F>
int guarded(F &&f) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors are returned, not printed
  try {
    f();
    return ELP_OK;
  } catch (const H5Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "out of host memory";
    return ELP_EINVAL;
  } catch (const std::exception &e) {
    g_err = e.what();
    return ELP_EINVAL;
  }
}

}  // namespace

extern "C" {

const char *elp_h5_last_error(void) { return g_err.c_str(); }

int elp_h5_read_analysis(const char *path, elp_h5_analysis *out) {
  return guarded([&] {
    if (!out) fail("null output");
    Hid f = open_file(path, H5F_ACC_RDONLY);
    read_analysis(f, out);
  });
}

int elp_h5_read_model(elp_model *m, const char *path, elp_h5_analysis *analysis, elp_h5_counts *counts) {
  return guarded([&] {
    if (!m) fail("null model");
    Hid f = open_file(path, H5F_ACC_RDONLY);
    elp_h5_analysis a;
    read_analysis(f, &a);
    elp_h5_counts c;
    std::memset(&c, 0, sizeof(c));
    read_model(m, f, &c);
    if (analysis) *analysis = a;
    if (counts) *counts = c;
  });
}

int elp_h5_write_solution(const char *path, int64_t n, int64_t n_freq, const double *freqs_hz, const double *X,
                          const int64_t *node_ids, const int32_t *dofs) {
  return guarded([&] {
    if (n <= 0 || n_freq < 0 || (n_freq > 0 && (!freqs_hz || !X)) || !node_ids || !dofs) fail("invalid arguments");
    Hid f = create_or_open(path, true);
    Hid sol = ensure_group(f, "/Solution");
    Hid state = ensure_group(f, "/Solution/State");
    Hid ct = complex_type();
    const hsize_t dim = (hsize_t)n;
    for (int64_t k = 0; k < n_freq; k++) {
      const std::string step = std::to_string(k + 1);
      write_dataset(state, "vecFemStep" + step, ct, 1, &dim, X + 2 * n * k);
      write_string_attr(state, "FreqStep" + step, std::to_string(freqs_hz[k]));
    }
    Hid maps = ensure_group(f, "/Solution/Maps");
    std::vector<int32_t> map(2 * n);
    for (int64_t r = 0; r < n; r++) {
      map[2 * r] = to_i32(node_ids[r], "a node id");
      map[2 * r + 1] = dofs[r];
    }
    const hsize_t dims[2] = {(hsize_t)n, 2};
    write_dataset(maps, "mtxDofMap", H5T_NATIVE_INT, 2, dims, map.data());
  });
}

int elp_h5_write_csr(const char *path, const char *name, int64_t n, const int32_t *rowptr, const int32_t *colind,
                     const double *vals) {
  return guarded([&] {
    if (!name || !*name || n <= 0 || !rowptr || !colind || !vals) fail("invalid arguments");
    if (std::strchr(name, '/')) fail("matrix name must not contain '/'");
    const int64_t nnz = rowptr[n];
    if (rowptr[0] != 0 || nnz < 0) fail("rowptr must start at 0");
    for (int64_t r = 0; r < n; r++)
      if (rowptr[r + 1] < rowptr[r]) fail("rowptr decreases at row " + std::to_string(r));
    for (int64_t k = 0; k < nnz; k++)
      if (colind[k] < 0 || colind[k] >= n) fail("column index out of range at entry " + std::to_string(k));
    Hid f = create_or_open(path, false);
    Hid top = ensure_group(f, "/SystemMatrices");
    const std::string gname = std::string("/SystemMatrices/") + name;
    if (H5Lexists(f, gname.c_str(), H5P_DEFAULT) > 0) fail("group " + gname + " exists already");
    Hid g = ensure_group(f, gname);
    write_int_attr(g, "numrow", to_i32(n, "the row count"));
    write_int_attr(g, "numcol", to_i32(n, "the column count"));
    std::vector<int32_t> ia(n + 1), ja(nnz);
    for (int64_t r = 0; r <= n; r++) ia[r] = rowptr[r] + 1;  // 1-based (mathlibrary.cpp:62)
    for (int64_t k = 0; k < nnz; k++) ja[k] = colind[k] + 1;
    hsize_t d1 = (hsize_t)(n + 1), d2 = (hsize_t)nnz;
    write_dataset(g, "vecCsrIa", H5T_NATIVE_INT, 1, &d1, ia.data());
    write_dataset(g, "vecCsrJa", H5T_NATIVE_INT, 1, &d2, ja.data());
    Hid ct = complex_type();
    write_dataset(g, "cmpCsrVal", ct, 1, &d2, vals);
  });
}

int elp_h5_read_complex_vector(const char *path, const char *dataset, int64_t n, double *X) {
  return guarded([&] {
    if (!dataset || n < 0 || (n > 0 && !X)) fail("invalid arguments");
    Hid f = open_file(path, H5F_ACC_RDONLY);
    if (!exists(f, dataset)) fail(std::string("missing dataset ") + dataset);
    Hid d(H5Dopen2(f, dataset, H5P_DEFAULT), H5Dclose);
    if (d < 0) fail(std::string("cannot open ") + dataset);
    Hid sp(H5Dget_space(d), H5Sclose);
    if (H5Sget_simple_extent_npoints(sp) != n) fail(std::string(dataset) + ": length differs from n");
    Hid ct = complex_type();
    if (n > 0 && H5Dread(d, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, X) < 0)
      fail(std::string(dataset) + ": not a compound {real, imag} vector");
  });
}

}  // extern "C"
