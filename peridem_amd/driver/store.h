// store.h -- dataset containers of the host driver.
//
// The reference reads and writes HDF5 with 2-D datasets (rw_hdf5.h:53-265:
// store_rowvec / store_col / load_rowvecs / load_col / load_rank1).  Store
// keeps that model -- named datasets with a type and dims, groups implied by
// '/' in the path -- behind two containers:
//   H5Store    HDF5 files (built when PD_WITH_HDF5 is defined)
//   BinStore   "pdbin": a flat binary file of the same named datasets, so the
//              driver and its tests do not depend on an HDF5 installation.
// Integer datasets are read as uint32/int32 whatever their stored integer
// type, as the reference's NATIVE_UINT reads do.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef PD_WITH_HDF5
#include <H5Cpp.h>
#endif

struct DSet {
  char type = 'd';                 // 'd' f64, 'u' u32, 'i' i32
  std::vector<uint64_t> dims;
  std::vector<uint8_t> bytes;
};

class Store {
 public:
  virtual ~Store() {}
  virtual bool has(const std::string& path) = 0;
  virtual DSet get(const std::string& path, char as) = 0;
  virtual void put(const std::string& path, const DSet& d) = 0;
  virtual std::vector<std::string> names() = 0;
  virtual char stored_type(const std::string& path) = 0;

  std::vector<double> read_f64(const std::string& p) { return as_vec<double>(get(p, 'd')); }
  std::vector<uint32_t> read_u32(const std::string& p) { return as_vec<uint32_t>(get(p, 'u')); }
  std::vector<int32_t> read_i32(const std::string& p) { return as_vec<int32_t>(get(p, 'i')); }
  void write_f64(const std::string& p, const std::vector<double>& v, std::vector<uint64_t> dims) {
    put(p, make(v, 'd', dims));
  }
  void write_u32(const std::string& p, const std::vector<uint32_t>& v, std::vector<uint64_t> dims) {
    put(p, make(v, 'u', dims));
  }
  void write_i32(const std::string& p, const std::vector<int32_t>& v, std::vector<uint64_t> dims) {
    put(p, make(v, 'i', dims));
  }

 protected:
  template <typename T>
  static std::vector<T> as_vec(const DSet& d) {
    std::vector<T> v(d.bytes.size() / sizeof(T));
    if (!v.empty()) memcpy(v.data(), d.bytes.data(), d.bytes.size());
    return v;
  }
  template <typename T>
  static DSet make(const std::vector<T>& v, char t, const std::vector<uint64_t>& dims) {
    DSet d;
    d.type = t;
    d.dims = dims;
    d.bytes.resize(v.size() * sizeof(T));
    if (!v.empty()) memcpy(d.bytes.data(), v.data(), d.bytes.size());
    return d;
  }
};

// pdbin: "PDBIN001" then records {u32 name length, name, char type, u32 ndims,
// u64 dims[ndims], raw little-endian data}.
class BinStore : public Store {
 public:
  BinStore(const std::string& fn, bool write) : fn_(fn), write_(write) {
    if (write) return;
    FILE* f = fopen(fn.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + fn);
    char magic[8];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "PDBIN001", 8)) {
      fclose(f);
      throw std::runtime_error(fn + " is not a pdbin file");
    }
    for (;;) {
      uint32_t nl;
      if (fread(&nl, 4, 1, f) != 1) break;
      std::string name(nl, '\0');
      DSet d;
      uint32_t nd;
      bool ok = fread(&name[0], 1, nl, f) == nl && fread(&d.type, 1, 1, f) == 1 && fread(&nd, 4, 1, f) == 1;
      d.dims.resize(nd);
      ok = ok && fread(d.dims.data(), 8, nd, f) == nd;
      uint64_t n = 1;
      for (auto x : d.dims) n *= x;
      size_t es = d.type == 'd' ? 8 : 4;
      d.bytes.resize(n * es);
      ok = ok && (n == 0 || fread(d.bytes.data(), 1, n * es, f) == n * es);
      if (!ok) {
        fclose(f);
        throw std::runtime_error(fn + ": truncated record");
      }
      m_[name] = std::move(d);
    }
    fclose(f);
  }
  ~BinStore() override {
    if (!write_) return;
    FILE* f = fopen(fn_.c_str(), "wb");
    if (!f) return;
    fwrite("PDBIN001", 1, 8, f);
    for (auto& kv : m_) {
      uint32_t nl = (uint32_t)kv.first.size(), nd = (uint32_t)kv.second.dims.size();
      fwrite(&nl, 4, 1, f);
      fwrite(kv.first.data(), 1, nl, f);
      fwrite(&kv.second.type, 1, 1, f);
      fwrite(&nd, 4, 1, f);
      fwrite(kv.second.dims.data(), 8, nd, f);
      if (!kv.second.bytes.empty()) fwrite(kv.second.bytes.data(), 1, kv.second.bytes.size(), f);
    }
    fclose(f);
  }
  bool has(const std::string& p) override { return m_.count(strip(p)) != 0; }
  DSet get(const std::string& p, char as) override {
    auto it = m_.find(strip(p));
    if (it == m_.end()) throw std::runtime_error(fn_ + ": missing dataset " + p);
    const DSet& d = it->second;
    if (d.type == as) return d;
    // numeric conversion between the stored and the requested type
    DSet o;
    o.type = as;
    o.dims = d.dims;
    size_t n = d.bytes.size() / (d.type == 'd' ? 8 : 4);
    o.bytes.resize(n * (as == 'd' ? 8 : 4));
    for (size_t i = 0; i < n; i++) {
      double v;
      if (d.type == 'd') memcpy(&v, &d.bytes[8 * i], 8);
      else if (d.type == 'u') { uint32_t x; memcpy(&x, &d.bytes[4 * i], 4); v = x; }
      else { int32_t x; memcpy(&x, &d.bytes[4 * i], 4); v = x; }
      if (as == 'd') memcpy(&o.bytes[8 * i], &v, 8);
      else if (as == 'u') { uint32_t x = (uint32_t)v; memcpy(&o.bytes[4 * i], &x, 4); }
      else { int32_t x = (int32_t)v; memcpy(&o.bytes[4 * i], &x, 4); }
    }
    return o;
  }
  void put(const std::string& p, const DSet& d) override { m_[strip(p)] = d; }
  char stored_type(const std::string& p) override {
    auto it = m_.find(strip(p));
    if (it == m_.end()) throw std::runtime_error(fn_ + ": missing dataset " + p);
    return it->second.type;
  }
  std::vector<std::string> names() override {
    std::vector<std::string> v;
    for (auto& kv : m_) v.push_back(kv.first);
    return v;
  }

 private:
  static std::string strip(const std::string& p) { return (!p.empty() && p[0] == '/') ? p.substr(1) : p; }
  std::string fn_;
  bool write_;
  std::map<std::string, DSet> m_;
};

#ifdef PD_WITH_HDF5
class H5Store : public Store {
 public:
  H5Store(const std::string& fn, bool write)
      : f_(fn.c_str(), write ? H5F_ACC_TRUNC : H5F_ACC_RDONLY), fn_(fn) {}
  bool has(const std::string& p) override {
    try {
      H5::Exception::dontPrint();
      return H5Lexists(f_.getId(), p.c_str(), H5P_DEFAULT) > 0;
    } catch (...) {
      return false;
    }
  }
  DSet get(const std::string& p, char as) override {
    H5::DataSet ds;
    try {
      ds = f_.openDataSet(p.c_str());
    } catch (H5::Exception&) {
      throw std::runtime_error(fn_ + ": missing dataset " + p);
    }
    H5::DataSpace sp = ds.getSpace();
    int nd = sp.getSimpleExtentNdims();
    std::vector<hsize_t> dims(nd > 0 ? nd : 1, 1);
    if (nd > 0) sp.getSimpleExtentDims(dims.data());
    DSet d;
    d.type = as;
    uint64_t n = 1;
    for (int k = 0; k < nd; k++) {
      d.dims.push_back(dims[k]);
      n *= dims[k];
    }
    d.bytes.resize(n * (as == 'd' ? 8 : 4));
    if (n) {
      const H5::PredType& t = as == 'd' ? H5::PredType::NATIVE_DOUBLE
                                        : (as == 'u' ? H5::PredType::NATIVE_UINT : H5::PredType::NATIVE_INT);
      ds.read(d.bytes.data(), t);
    }
    return d;
  }
  void put(const std::string& p, const DSet& d) override {
    // create the groups on the path
    for (size_t i = 1; i < p.size(); i++)
      if (p[i] == '/') {
        std::string g = p.substr(0, i);
        if (H5Lexists(f_.getId(), g.c_str(), H5P_DEFAULT) <= 0) f_.createGroup(g.c_str());
      }
    std::vector<hsize_t> dims(d.dims.begin(), d.dims.end());
    H5::DataSpace sp((int)dims.size(), dims.data());
    const H5::PredType& t = d.type == 'd' ? H5::PredType::NATIVE_DOUBLE
                                          : (d.type == 'u' ? H5::PredType::NATIVE_UINT : H5::PredType::NATIVE_INT);
    H5::DataSet ds = f_.createDataSet(p.c_str(), t, sp);
    if (!d.bytes.empty()) ds.write(d.bytes.data(), t);
  }
  std::vector<std::string> names() override {
    std::vector<std::string> out;
    walk("/", out);
    return out;
  }
  char stored_type(const std::string& p) override {
    H5::DataSet ds = f_.openDataSet(p.c_str());
    if (ds.getTypeClass() == H5T_FLOAT) return 'd';
    return ds.getIntType().getSign() == H5T_SGN_NONE ? 'u' : 'i';
  }

 private:
  void walk(const std::string& g, std::vector<std::string>& out) {
    H5::Group grp = f_.openGroup(g.c_str());
    for (hsize_t i = 0; i < grp.getNumObjs(); i++) {
      std::string nm = grp.getObjnameByIdx(i);
      std::string full = (g == "/" ? "" : g) + "/" + nm;
      if (grp.getObjTypeByIdx(i) == H5G_GROUP) walk(full, out);
      else out.push_back(full.substr(1));
    }
  }
  H5::H5File f_;
  std::string fn_;
};
#endif

inline std::unique_ptr<Store> open_store(const std::string& fn, bool write) {
  bool h5 = fn.size() > 3 && fn.compare(fn.size() - 3, 3, ".h5") == 0;
#ifdef PD_WITH_HDF5
  if (h5) return std::unique_ptr<Store>(new H5Store(fn, write));
#else
  if (h5) {
    // without HDF5 the same name holds a pdbin container
    return std::unique_ptr<Store>(new BinStore(fn, write));
  }
#endif
  return std::unique_ptr<Store>(new BinStore(fn, write));
}

inline void copy_store(Store& a, Store& b) {
  for (auto& n : a.names()) b.put(n, a.get(n, a.stored_type(n)));
}
