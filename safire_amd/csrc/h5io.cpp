// Readers for the reference's HDF5 inputs (declared in include/afq_h5.h), host code on the
// HDF5 1.10 C API.  Layouts follow RealDenseHamiltonian.cpp:76-136, hdf5_helpers.hpp:84-108,
// readWfn.cpp:67-110 and 148-214, readWfn.h:90-156 and csr_utils.hpp:353-420.
#include "afq_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
constexpr int kErrArg = -1, kErrShape = -3;

// RAII handle with the matching H5*close
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id() = default;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  Id(Id&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

Id open_file(const char* path) {
  if (!path) throw Fail(kErrArg, "null path");
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw Fail(kErrArg, std::string("cannot open HDF5 file ") + path);
  return Id(f, H5Fclose);
}

// every component of a slash-separated path exists below loc
bool exists(hid_t loc, const std::string& path) {
  std::string acc;
  size_t p = 0;
  while (p <= path.size()) {
    size_t q = path.find('/', p);
    if (q == std::string::npos) q = path.size();
    if (q > p) {
      acc += (acc.empty() ? "" : "/") + path.substr(p, q - p);
      if (H5Lexists(loc, acc.c_str(), H5P_DEFAULT) <= 0) return false;
    }
    p = q + 1;
  }
  return true;
}

Id open_group(hid_t loc, const std::string& path) {
  if (!exists(loc, path)) throw Fail(kErrArg, "missing group " + path);
  hid_t g = H5Gopen2(loc, path.c_str(), H5P_DEFAULT);
  if (g < 0) throw Fail(kErrArg, "cannot open group " + path);
  return Id(g, H5Gclose);
}

Id open_dataset(hid_t loc, const std::string& path) {
  if (!exists(loc, path)) throw Fail(kErrArg, "missing dataset " + path);
  hid_t d = H5Dopen2(loc, path.c_str(), H5P_DEFAULT);
  if (d < 0) throw Fail(kErrArg, "cannot open dataset " + path);
  return Id(d, H5Dclose);
}

std::vector<hsize_t> dims_of(hid_t dset) {
  Id sp(H5Dget_space(dset), H5Sclose);
  int nd = H5Sget_simple_extent_ndims(sp);
  if (nd < 0) throw Fail(kErrArg, "bad dataspace");
  std::vector<hsize_t> d(nd);
  if (nd > 0) H5Sget_simple_extent_dims(sp, d.data(), nullptr);
  return d;
}

hsize_t count_of(const std::vector<hsize_t>& d) {
  hsize_t n = 1;
  for (hsize_t x : d) n *= x;
  return n;
}

std::vector<double> read_doubles(hid_t loc, const std::string& path, std::vector<hsize_t>* dims) {
  Id d = open_dataset(loc, path);
  auto dd = dims_of(d);
  std::vector<double> v(count_of(dd));
  if (!v.empty() && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    throw Fail(kErrArg, "cannot read " + path + " as double");
  if (dims) *dims = dd;
  return v;
}

std::vector<int> read_ints(hid_t loc, const std::string& path, std::vector<hsize_t>* dims) {
  Id d = open_dataset(loc, path);
  auto dd = dims_of(d);
  std::vector<int> v(count_of(dd));
  if (!v.empty() && H5Dread(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    throw Fail(kErrArg, "cannot read " + path + " as int");
  if (dims) *dims = dd;
  return v;
}

struct C128 {
  double r, i;
};

// complex values: the reference writes compound {r, i}; a real dataset reads with zero imaginary
std::vector<C128> read_complex(hid_t loc, const std::string& path, std::vector<hsize_t>* dims) {
  Id d = open_dataset(loc, path);
  auto dd = dims_of(d);
  std::vector<C128> v(count_of(dd));
  Id ft(H5Dget_type(d), H5Tclose);
  if (H5Tget_class(ft) == H5T_COMPOUND) {
    Id mt(H5Tcreate(H5T_COMPOUND, sizeof(C128)), H5Tclose);
    H5Tinsert(mt, "r", HOFFSET(C128, r), H5T_NATIVE_DOUBLE);
    H5Tinsert(mt, "i", HOFFSET(C128, i), H5T_NATIVE_DOUBLE);
    if (!v.empty() && H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
      throw Fail(kErrArg, "cannot read " + path + " as complex {r, i}");
  } else {
    std::vector<double> re(v.size());
    if (!re.empty() && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, re.data()) < 0)
      throw Fail(kErrArg, "cannot read " + path);
    for (size_t k = 0; k < re.size(); ++k) v[k] = {re[k], 0.0};
  }
  if (dims) *dims = dd;
  return v;
}

bool attr_exists(hid_t obj, const char* name) { return H5Aexists(obj, name) > 0; }

double attr_double(hid_t obj, const char* name) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (a < 0) throw Fail(kErrArg, std::string("missing attribute ") + name);
  double v = 0;
  if (H5Aread(a, H5T_NATIVE_DOUBLE, &v) < 0) throw Fail(kErrArg, std::string("cannot read ") + name);
  return v;
}

int attr_int(hid_t obj, const char* name) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (a < 0) throw Fail(kErrArg, std::string("missing attribute ") + name);
  int v = 0;
  if (H5Aread(a, H5T_NATIVE_INT, &v) < 0) throw Fail(kErrArg, std::string("cannot read ") + name);
  return v;
}

std::string attr_string(hid_t obj, const char* name) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (a < 0) throw Fail(kErrArg, std::string("missing attribute ") + name);
  Id ft(H5Aget_type(a), H5Tclose);
  if (H5Tget_class(ft) != H5T_STRING) throw Fail(kErrArg, std::string(name) + " is not a string");
  std::string out;
  if (H5Tis_variable_str(ft) > 0) {
    Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, H5T_VARIABLE);
    H5Tset_cset(mt, H5Tget_cset(ft));
    char* s = nullptr;
    if (H5Aread(a, mt, &s) < 0) throw Fail(kErrArg, std::string("cannot read ") + name);
    if (s) {
      out = s;
      H5free_memory(s);
    }
  } else {
    size_t n = H5Tget_size(ft);
    std::vector<char> buf(n + 1, 0);
    Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, n + 1);
    H5Tset_cset(mt, H5Tget_cset(ft));
    if (H5Aread(a, mt, buf.data()) < 0) throw Fail(kErrArg, std::string("cannot read ") + name);
    out = buf.data();
  }
  return out;
}

// walker type of a spin_type string (WALKER_TYPES names, case-insensitive)
int walker_type_of(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  if (s == "closed") return AFQ_H5_CLOSED;
  if (s == "collinear") return AFQ_H5_COLLINEAR;
  if (s == "noncollinear") return AFQ_H5_NONCOLLINEAR;
  throw Fail(kErrArg, "unknown spin_type \"" + s + "\"");
}

const char* kind_name(int kind) {
  if (kind == AFQ_H5_NOMSD) return "NOMSD";
  if (kind == AFQ_H5_PHMSD) return "PHMSD";
  throw Fail(kErrArg, "unknown wavefunction kind");
}

// (rows, cols) of a sparse matrix group; legacy groups store dims = (rows, cols, nnz)
void csr_shape(hid_t g, long& rows, long& cols, bool& legacy) {
  legacy = !exists(g, "shape");
  auto d = read_ints(g, legacy ? "dims" : "shape", nullptr);
  if (d.size() != (legacy ? 3u : 2u)) throw Fail(kErrShape, "sparse matrix shape has a wrong length");
  rows = d[0];
  cols = d[1];
}

void csr_to_dense(hid_t g, long rows, long cols, double* out) {
  long r0, c0;
  bool legacy;
  csr_shape(g, r0, c0, legacy);
  if (r0 != rows || c0 != cols)
    throw Fail(kErrShape, "sparse matrix is " + std::to_string(r0) + " x " + std::to_string(c0) +
                              ", expected " + std::to_string(rows) + " x " + std::to_string(cols));
  std::vector<int> pb, pe;
  if (legacy) {
    pb = read_ints(g, "pointers_begin_", nullptr);
    pe = read_ints(g, "pointers_end_", nullptr);
  } else {
    auto rp = read_ints(g, "row_pointers", nullptr);
    if ((long)rp.size() != rows + 1) throw Fail(kErrShape, "row_pointers has a wrong length");
    pb.assign(rp.begin(), rp.end() - 1);
    pe.assign(rp.begin() + 1, rp.end());
  }
  if ((long)pb.size() != rows || (long)pe.size() != rows) throw Fail(kErrShape, "row pointers");
  auto col = read_ints(g, legacy ? "jdata_" : "column_indices", nullptr);
  auto val = read_complex(g, legacy ? "data_" : "values", nullptr);
  std::fill(out, out + 2 * rows * cols, 0.0);
  for (long r = 0; r < rows; ++r)
    for (int k = pb[r]; k < pe[r]; ++k) {
      if (k < 0 || k >= (int)col.size() || k >= (int)val.size())
        throw Fail(kErrShape, "sparse matrix pointer out of range");
      const int c = col[k];
      if (c < 0 || c >= cols) throw Fail(kErrShape, "sparse matrix column out of range");
      out[2 * (r * cols + c)] += val[k].r;
      out[2 * (r * cols + c) + 1] += val[k].i;
    }
}

int npsit_of(hid_t g) {
  int n = 0;
  while (exists(g, "PsiT_" + std::to_string(n))) ++n;
  return n;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Fail& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return kErrArg;
  }
}

// ---- results file ----------------------------------------------------------------------------

// group `name` below loc, created when missing
Id open_or_create_group(hid_t loc, const std::string& name) {
  if (name.empty()) throw Fail(kErrArg, "empty group name in an observable path");
  if (H5Lexists(loc, name.c_str(), H5P_DEFAULT) > 0) {
    hid_t g = H5Gopen2(loc, name.c_str(), H5P_DEFAULT);
    if (g < 0) throw Fail(kErrArg, "cannot open group " + name);
    return Id(g, H5Gclose);
  }
  hid_t g = H5Gcreate2(loc, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) throw Fail(kErrArg, "cannot create group " + name);
  return Id(g, H5Gclose);
}

std::vector<std::string> split_path(const std::string& path) {
  std::vector<std::string> out;
  size_t p = 0;
  while (p <= path.size()) {
    size_t q = path.find('/', p);
    if (q == std::string::npos) q = path.size();
    out.push_back(path.substr(p, q - p));
    p = q + 1;
  }
  return out;
}

std::string bins_path(const char* stage, const char* name) {
  if (!stage || !name || !*stage || !*name) throw Fail(kErrArg, "stage and observable name needed");
  return std::string("Measurements/") + stage + "/" + name + "/bins";
}

void write_string_attr(hid_t obj, const char* key, const char* value) {
  Id t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(t, std::strlen(value) + 1);
  H5Tset_strpad(t, H5T_STR_NULLTERM);
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(obj, key, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (a < 0 || H5Awrite(a, t, value) < 0) throw Fail(kErrArg, std::string("cannot write attribute ") + key);
}

struct NameCollector {
  std::vector<std::string> names;
};

herr_t collect_bins(hid_t, const char* name, const H5O_info_t* info, void* op) {
  if (info->type != H5O_TYPE_DATASET) return 0;
  std::string n(name);
  const std::string tail = "/bins";
  if (n.size() > tail.size() && n.compare(n.size() - tail.size(), tail.size(), tail) == 0)
    static_cast<NameCollector*>(op)->names.push_back(n.substr(0, n.size() - tail.size()));
  return 0;
}

}  // namespace

extern "C" {

int afq_h5_results_append(const char* path, const char* stage, const char* name, int rank,
                          const long* dims, int is_complex, long nbins, const double* data) {
  return guarded([&] {
    if (!path || rank < 0 || rank > 8 || (rank > 0 && !dims) || nbins < 0 || (nbins > 0 && !data))
      throw Fail(kErrArg, "invalid results arguments");
    const std::string bp = bins_path(stage, name);
    quiet();
    namespace fs = std::filesystem;
    const fs::path target(path), tmp = fs::path(std::string(path) + ".tmp");
    const bool existed = fs::exists(target);
    if (existed) fs::copy_file(target, tmp, fs::copy_options::overwrite_existing);
    {
      Id f(existed ? H5Fopen(tmp.c_str(), H5F_ACC_RDWR, H5P_DEFAULT)
                   : H5Fcreate(tmp.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
           H5Fclose);
      if (f < 0) throw Fail(kErrArg, std::string("cannot open ") + tmp.string() + " for writing");
      Id root(H5Gopen2(f, "/", H5P_DEFAULT), H5Gclose);
      if (!attr_exists(root, "format_version")) {
        const int v = 1;
        Id sp(H5Screate(H5S_SCALAR), H5Sclose);
        Id a(H5Acreate2(root, "format_version", H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT),
             H5Aclose);
        if (a < 0 || H5Awrite(a, H5T_NATIVE_INT, &v) < 0)
          throw Fail(kErrArg, "cannot write format_version");
      }
      // Measurements/<stage>/<name components>
      std::vector<std::string> parts = split_path(std::string("Measurements/") + stage + "/" + name);
      std::vector<Id> groups;
      groups.push_back(open_or_create_group(root, parts[0]));
      for (size_t i = 1; i < parts.size(); ++i) groups.push_back(open_or_create_group(groups.back(), parts[i]));
      hid_t g = groups.back();
      const int frank = rank + 1 + (is_complex ? 1 : 0);
      std::vector<hsize_t> want(frank);
      want[0] = (hsize_t)nbins;
      for (int i = 0; i < rank; ++i) {
        if (dims[i] < 0) throw Fail(kErrArg, "negative extent");
        want[1 + i] = (hsize_t)dims[i];
      }
      if (is_complex) want[frank - 1] = 2;
      if (H5Lexists(g, "bins", H5P_DEFAULT) <= 0) {
        std::vector<hsize_t> d0 = want, maxd = want, chunk = want;
        d0[0] = 0;
        maxd[0] = H5S_UNLIMITED;
        for (auto& c : chunk) c = std::max<hsize_t>(c, 1);
        Id cp(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
        if (H5Pset_chunk(cp, frank, chunk.data()) < 0) throw Fail(kErrArg, "H5Pset_chunk failed");
        H5Pset_deflate(cp, 1);
        Id sp(H5Screate_simple(frank, d0.data(), maxd.data()), H5Sclose);
        Id ds(H5Dcreate2(g, "bins", H5T_IEEE_F64LE, sp, H5P_DEFAULT, cp, H5P_DEFAULT), H5Dclose);
        if (ds < 0) throw Fail(kErrArg, "cannot create " + bp);
        if (is_complex) write_string_attr(ds, "__complex__", "1");
      }
      Id ds(H5Dopen2(g, "bins", H5P_DEFAULT), H5Dclose);
      if (ds < 0) throw Fail(kErrArg, "cannot open " + bp);
      std::vector<hsize_t> cur = dims_of(ds);
      const bool cplx = attr_exists(ds, "__complex__");
      if ((int)cur.size() != frank || cplx != (is_complex != 0))
        throw Fail(kErrShape, bp + " holds bins of another rank or type");
      for (int i = 1; i < frank; ++i)
        if (cur[i] != want[i]) throw Fail(kErrShape, bp + " holds bins of another shape");
      if (nbins > 0) {
        std::vector<hsize_t> grown = cur;
        grown[0] += (hsize_t)nbins;
        if (H5Dset_extent(ds, grown.data()) < 0) throw Fail(kErrArg, "H5Dset_extent failed");
        Id fs_(H5Dget_space(ds), H5Sclose);
        std::vector<hsize_t> start(frank, 0);
        start[0] = cur[0];
        if (H5Sselect_hyperslab(fs_, H5S_SELECT_SET, start.data(), nullptr, want.data(), nullptr) < 0)
          throw Fail(kErrArg, "cannot select the appended bins");
        Id ms(H5Screate_simple(frank, want.data(), nullptr), H5Sclose);
        if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, ms, fs_, H5P_DEFAULT, data) < 0)
          throw Fail(kErrArg, "cannot write " + bp);
      }
    }
    fs::rename(tmp, target);
  });
}

int afq_h5_results_shape(const char* path, const char* stage, const char* name, long* nbins,
                         int* rank, long* dims, int* is_complex) {
  return guarded([&] {
    if (!nbins || !rank || !dims || !is_complex) throw Fail(kErrArg, "null output");
    Id f = open_file(path);
    const std::string bp = bins_path(stage, name);
    if (!exists(f, bp)) throw Fail(kErrArg, "no " + bp + " in " + path);
    Id ds(H5Dopen2(f, bp.c_str(), H5P_DEFAULT), H5Dclose);
    std::vector<hsize_t> d = dims_of(ds);
    const bool cplx = attr_exists(ds, "__complex__");
    const int r = (int)d.size() - 1 - (cplx ? 1 : 0);
    if (r < 0 || r > 8) throw Fail(kErrShape, bp + " has an unexpected rank");
    *nbins = (long)d[0];
    *rank = r;
    for (int i = 0; i < r; ++i) dims[i] = (long)d[1 + i];
    *is_complex = cplx ? 1 : 0;
  });
}

int afq_h5_results_read(const char* path, const char* stage, const char* name, double* out) {
  return guarded([&] {
    if (!out) throw Fail(kErrArg, "null output");
    Id f = open_file(path);
    const std::string bp = bins_path(stage, name);
    if (!exists(f, bp)) throw Fail(kErrArg, "no " + bp + " in " + path);
    Id ds(H5Dopen2(f, bp.c_str(), H5P_DEFAULT), H5Dclose);
    if (count_of(dims_of(ds)) == 0) return;
    if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
      throw Fail(kErrArg, "cannot read " + bp);
  });
}

int afq_h5_results_names(const char* path, const char* stage, char* buf, long cap, long* len) {
  return guarded([&] {
    if (!len || !stage) throw Fail(kErrArg, "null argument");
    Id f = open_file(path);
    const std::string sp = std::string("Measurements/") + stage;
    if (!exists(f, sp)) throw Fail(kErrArg, "no " + sp + " in " + path);
    Id g(H5Gopen2(f, sp.c_str(), H5P_DEFAULT), H5Gclose);
    NameCollector nc;
    if (H5Ovisit(g, H5_INDEX_NAME, H5_ITER_INC, collect_bins, &nc) < 0)
      throw Fail(kErrArg, "cannot list " + sp);
    std::sort(nc.names.begin(), nc.names.end());
    std::string all;
    for (size_t i = 0; i < nc.names.size(); ++i) all += (i ? "\n" : "") + nc.names[i];
    *len = (long)all.size() + 1;
    if (buf && cap >= *len) std::memcpy(buf, all.c_str(), all.size() + 1);
  });
}

const char* afq_h5_last_error(void) { return g_err.c_str(); }

int afq_h5_hamiltonian_shape(const char* path, int* nspin, int* N, int* M, int* nspin_L) {
  return guarded([&] {
    Id f = open_file(path);
    Id hg = open_group(f, "Hamiltonian");
    std::vector<hsize_t> dh, dl;
    {
      Id d = open_dataset(hg, "hcore");
      dh = dims_of(d);
    }
    {
      Id d = open_dataset(hg, "DenseFactorized/L");
      dl = dims_of(d);
    }
    if (dh.size() != 5 || dh[1] != 1 || dh[3] != 1 || dh[2] != dh[4])
      throw Fail(kErrShape, "hcore must be [nspin][1][N][1][N]");
    if (dl.size() != 6 || dl[1] != 1 || dl[3] != 1 || dl[2] != dh[2] || dl[4] != dh[2])
      throw Fail(kErrShape, "DenseFactorized/L must be [nspin][1][N][1][N][M]");
    if (nspin) *nspin = (int)dh[0];
    if (N) *N = (int)dh[2];
    if (M) *M = (int)dl[5];
    if (nspin_L) *nspin_L = (int)dl[0];
  });
}

int afq_h5_read_hamiltonian(const char* path, int nelec, double* hcore, double* L, double* E0) {
  return guarded([&] {
    int nspin = 0, N = 0, M = 0, nsl = 0;
    if (afq_h5_hamiltonian_shape(path, &nspin, &N, &M, &nsl) != 0) throw Fail(kErrShape, g_err);
    // L is shared by every spin (nspin_L = 1) or has hcore's spin blocks
    // (RealDenseHamiltonian.cpp:98-111)
    if (nsl != 1 && nsl != nspin)
      throw Fail(kErrShape, "DenseFactorized/L must have 1 spin block or as many as hcore");
    Id f = open_file(path);
    Id hg = open_group(f, "Hamiltonian");
    if (hcore) {
      auto h = read_doubles(hg, "hcore", nullptr);
      std::copy(h.begin(), h.end(), hcore);
    }
    if (L) {
      auto l = read_doubles(hg, "DenseFactorized/L", nullptr);
      std::copy(l.begin(), l.end(), L);
    }
    if (E0) {
      double nuc = attr_exists(hg, "nuclear_energy") ? attr_double(hg, "nuclear_energy") : 0.0;
      double fzc = attr_exists(hg, "frozen_core_energy") ? attr_double(hg, "frozen_core_energy") : 0.0;
      double mad = attr_exists(hg, "madelung_constant") ? attr_double(hg, "madelung_constant") : 0.0;
      *E0 = nuc + fzc - mad * (double)nelec;
    }
  });
}

int afq_h5_wavefunction_info(const char* path, int* kind, int* walker_type, int* N, int* nup,
                             int* ndown, int* ndets, int* npsit) {
  return guarded([&] {
    Id f = open_file(path);
    const bool ph = exists(f, "Wavefunction/PHMSD");
    if (!ph && !exists(f, "Wavefunction/NOMSD"))
      throw Fail(kErrArg, "no /Wavefunction/NOMSD or /Wavefunction/PHMSD group");
    Id g = open_group(f, ph ? "Wavefunction/PHMSD" : "Wavefunction/NOMSD");
    const int wt = walker_type_of(attr_string(g, "spin_type"));
    const int npol = wt == AFQ_H5_NONCOLLINEAR ? 2 : 1;
    std::vector<hsize_t> dc;
    { Id d = open_dataset(g, "ci_coeffs"); dc = dims_of(d); }
    int n = 0, nu = 0, nd = 0;
    if (ph) {
      n = attr_int(g, "number_of_orbitals");
      std::vector<hsize_t> da, db;
      { Id d = open_dataset(g, "occa"); da = dims_of(d); }
      { Id d = open_dataset(g, "occb"); db = dims_of(d); }
      if (da.size() != 2 || db.size() != 2) throw Fail(kErrShape, "occa/occb must be 2-d");
      nu = (int)da[1];
      nd = (int)db[1];
    } else {
      long r, c;
      bool legacy;
      Id p0 = open_group(g, "PsiT_0");
      csr_shape(p0, r, c, legacy);
      n = (int)(c / npol);
      nu = (int)r;
      if (wt == AFQ_H5_COLLINEAR) {
        Id p1 = open_group(g, "PsiT_1");
        long r1, c1;
        csr_shape(p1, r1, c1, legacy);
        nd = (int)r1;
      } else if (wt == AFQ_H5_CLOSED) {
        nd = nu;
      }
    }
    if (kind) *kind = ph ? AFQ_H5_PHMSD : AFQ_H5_NOMSD;
    if (walker_type) *walker_type = wt;
    if (N) *N = n;
    if (nup) *nup = nu;
    if (ndown) *ndown = nd;
    if (ndets) *ndets = dc.empty() ? 0 : (int)dc[0];
    if (npsit) *npsit = npsit_of(g);
  });
}

int afq_h5_read_psit(const char* path, int kind, int k, int rows, int cols, double* out) {
  return guarded([&] {
    if (!out || rows < 0 || cols < 0) throw Fail(kErrArg, "bad output");
    Id f = open_file(path);
    Id g = open_group(f, std::string("Wavefunction/") + kind_name(kind) + "/PsiT_" + std::to_string(k));
    csr_to_dense(g, rows, cols, out);
  });
}

int afq_h5_read_ci(const char* path, int kind, int ndets, double* ci) {
  return guarded([&] {
    Id f = open_file(path);
    Id g = open_group(f, std::string("Wavefunction/") + kind_name(kind));
    auto c = read_complex(g, "ci_coeffs", nullptr);
    if (ndets < 0 || ndets > (int)c.size()) throw Fail(kErrShape, "ci_coeffs has fewer entries");
    for (int d = 0; d < ndets; ++d) {
      ci[2 * d] = c[d].r;
      ci[2 * d + 1] = c[d].i;
    }
  });
}

int afq_h5_read_occupations(const char* path, int ndets, int nup, int ndown, int* occa, int* occb) {
  return guarded([&] {
    Id f = open_file(path);
    Id g = open_group(f, "Wavefunction/PHMSD");
    std::vector<hsize_t> da, db;
    auto a = read_ints(g, "occa", &da);
    auto b = read_ints(g, "occb", &db);
    if (da.size() != 2 || db.size() != 2 || (int)da[1] != nup || (int)db[1] != ndown ||
        (long)da[0] < ndets || (long)db[0] < ndets)
      throw Fail(kErrShape, "occa/occb do not match ndets x (nup, ndown)");
    std::copy(a.begin(), a.begin() + (long)ndets * nup, occa);
    std::copy(b.begin(), b.begin() + (long)ndets * ndown, occb);
  });
}

}  // extern "C"
