// NetCDF classic (64-bit offset) reader and writer; see lpm_ncio.hpp.
#include "lpm_ncio.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>

namespace lpm {
namespace nc {
namespace {

constexpr uint32_t kTagDim = 0x0A, kTagVar = 0x0B, kTagAtt = 0x0C;

int type_size(Type t) { return t == kDouble ? 8 : t == kInt ? 4 : 1; }
int64_t pad4(int64_t n) { return (n + 3) & ~int64_t(3); }

// big-endian encoding into a byte buffer
struct Out {
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int k = 3; k >= 0; --k) b.push_back((uint8_t)(v >> (8 * k)));
  }
  void u64(uint64_t v) {
    for (int k = 7; k >= 0; --k) b.push_back((uint8_t)(v >> (8 * k)));
  }
  void bytes(const void* p, size_t n) {
    const uint8_t* c = (const uint8_t*)p;
    b.insert(b.end(), c, c + n);
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string& s) {
    u32((uint32_t)s.size());
    bytes(s.data(), s.size());
  }
  void att(const Att& a) {
    name(a.name);
    u32((uint32_t)a.type);
    if (a.type == kChar) {
      u32((uint32_t)a.text.size());
      bytes(a.text.data(), a.text.size());
    } else if (a.type == kInt) {
      u32((uint32_t)a.ints.size());
      for (int32_t v : a.ints) u32((uint32_t)v);
    } else {
      u32((uint32_t)a.doubles.size());
      for (double v : a.doubles) {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        u64(u);
      }
    }
  }
  void atts(const std::vector<Att>& as) {
    if (as.empty()) {
      u32(0);
      u32(0);
      return;
    }
    u32(kTagAtt);
    u32((uint32_t)as.size());
    for (const Att& a : as) att(a);
  }
};

// big-endian decoding with bounds checks
struct In {
  const std::vector<uint8_t>& b;
  size_t p = 0;
  bool bad = false;
  explicit In(const std::vector<uint8_t>& buf) : b(buf) {}
  uint32_t u32() {
    if (p + 4 > b.size()) {
      bad = true;
      return 0;
    }
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) v = (v << 8) | b[p++];
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return (hi << 32) | u32();
  }
  std::string str(size_t n) {
    if (p + n > b.size()) {
      bad = true;
      return std::string();
    }
    std::string s((const char*)&b[p], n);
    p = (size_t)pad4((int64_t)(p + n));
    return s;
  }
  std::string name() { return str(u32()); }
  Att att() {
    Att a;
    a.name = name();
    a.type = (Type)u32();
    const uint32_t n = u32();
    if (bad || n > (1u << 28)) {
      bad = true;
      return a;
    }
    if (a.type == kChar) {
      a.text = str(n);
    } else if (a.type == kInt) {
      for (uint32_t k = 0; k < n && !bad; ++k) a.ints.push_back((int32_t)u32());
    } else if (a.type == kDouble) {
      for (uint32_t k = 0; k < n && !bad; ++k) {
        const uint64_t u = u64();
        double v;
        std::memcpy(&v, &u, 8);
        a.doubles.push_back(v);
      }
    } else {
      // other types: skip their values
      const int sz = a.type == 3 ? 2 : a.type == 5 ? 4 : 1;
      str((size_t)n * sz);
    }
    return a;
  }
  std::vector<Att> atts() {
    std::vector<Att> as;
    const uint32_t tag = u32(), n = u32();
    if (tag == 0 && n == 0) return as;
    if (tag != kTagAtt || n > (1u << 20)) {
      bad = true;
      return as;
    }
    for (uint32_t k = 0; k < n && !bad; ++k) as.push_back(att());
    return as;
  }
};

}  // namespace

int Writer::dim(const std::string& name, int64_t len) {
  dims_.push_back({name, len});
  return (int)dims_.size() - 1;
}

void Writer::att_text(const std::string& name, const std::string& v) {
  Att a;
  a.name = name;
  a.type = kChar;
  a.text = v;
  gatts_.push_back(a);
}

void Writer::att_int(const std::string& name, int32_t v) {
  Att a;
  a.name = name;
  a.type = kInt;
  a.ints = {v};
  gatts_.push_back(a);
}

void Writer::att_double(const std::string& name, double v) {
  Att a;
  a.name = name;
  a.type = kDouble;
  a.doubles = {v};
  gatts_.push_back(a);
}

void Writer::var(const std::string& name, Type type, const std::vector<int>& dims, const void* data,
                 const char* units) {
  VarInfo v;
  v.name = name;
  v.type = type;
  v.dims = dims;
  v.count = 1;
  for (int d : dims) v.count *= dims_[d].second;
  if (units) {
    Att a;
    a.name = "units";
    a.type = kChar;
    a.text = units;
    v.atts.push_back(a);
  }
  vars_.push_back(v);
  data_.push_back(data);
}

void Writer::var_copy(const std::string& name, Type type, const std::vector<int>& dims, const void* data,
                      const char* units) {
  int64_t count = 1;
  for (int d : dims) count *= dims_[d].second;
  const uint8_t* p = (const uint8_t*)data;
  owned_.emplace_back(p, p + count * type_size(type));
  var(name, type, dims, owned_.back().data(), units);
}

bool Writer::write(const std::string& path, std::string& err) const {
  // a fixed dimension of length 0 would read as the record dimension
  for (const auto& d : dims_)
    if (d.second <= 0 || d.second > 0x7FFFFFFF) {
      err = "dimension " + d.first + " has length " + std::to_string(d.second) + " (must be in [1, 2^31))";
      return false;
    }
  // the header twice: once to size it (begin offsets are fixed width)
  std::vector<int64_t> begins(vars_.size(), 0);
  Out h;
  for (int pass = 0; pass < 2; ++pass) {
    h.b.clear();
    h.bytes("CDF\x02", 4);
    h.u32(0);  // numrecs: no record variables
    if (dims_.empty()) {
      h.u32(0);
      h.u32(0);
    } else {
      h.u32(kTagDim);
      h.u32((uint32_t)dims_.size());
      for (const auto& d : dims_) {
        h.name(d.first);
        h.u32((uint32_t)d.second);
      }
    }
    h.atts(gatts_);
    if (vars_.empty()) {
      h.u32(0);
      h.u32(0);
    } else {
      h.u32(kTagVar);
      h.u32((uint32_t)vars_.size());
      for (size_t k = 0; k < vars_.size(); ++k) {
        const VarInfo& v = vars_[k];
        h.name(v.name);
        h.u32((uint32_t)v.dims.size());
        for (int d : v.dims) h.u32((uint32_t)d);
        h.atts(v.atts);
        h.u32((uint32_t)v.type);
        const int64_t vsize = pad4(v.count * type_size(v.type));
        h.u32(vsize >= 0xFFFFFFFCll ? 0xFFFFFFFFu : (uint32_t)vsize);
        h.u64((uint64_t)begins[k]);
      }
    }
    int64_t off = (int64_t)h.b.size();
    for (size_t k = 0; k < vars_.size(); ++k) {
      begins[k] = off;
      off += pad4(vars_[k].count * type_size(vars_[k].type));
    }
  }
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    err = "cannot open " + path + " for writing";
    return false;
  }
  f.write((const char*)h.b.data(), (std::streamsize)h.b.size());
  std::vector<uint8_t> buf;
  for (size_t k = 0; k < vars_.size(); ++k) {
    const VarInfo& v = vars_[k];
    const int sz = type_size(v.type);
    buf.assign((size_t)pad4(v.count * sz), 0);
    const uint8_t* src = (const uint8_t*)data_[k];
    for (int64_t i = 0; i < v.count; ++i)
      for (int c = 0; c < sz; ++c) buf[(size_t)(i * sz + c)] = src[i * sz + (sz - 1 - c)];  // to big-endian
    f.write((const char*)buf.data(), (std::streamsize)buf.size());
  }
  if (!f) {
    err = "write failed: " + path;
    return false;
  }
  return true;
}

bool Reader::open(const std::string& path, std::string& err) {
  path_ = path;
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  // the header lies at the front; read a generous prefix and grow if needed
  std::vector<uint8_t> b;
  for (size_t want = 1 << 16;; want *= 4) {
    b.assign(want, 0);
    f.clear();
    f.seekg(0);
    f.read((char*)b.data(), (std::streamsize)want);
    b.resize((size_t)f.gcount());
    In in(b);
    const std::string magic = in.str(4);
    if (in.bad || magic.substr(0, 3) != "CDF" || (magic[3] != 1 && magic[3] != 2)) {
      err = path + ": not a netCDF classic file";
      return false;
    }
    const bool offset64 = magic[3] == 2;
    in.u32();  // numrecs
    dims_.clear();
    gatts_.clear();
    vars_.clear();
    uint32_t tag = in.u32(), n = in.u32();
    if (tag == kTagDim)
      for (uint32_t k = 0; k < n && !in.bad; ++k) {
        std::string nm = in.name();
        dims_.push_back({nm, (int64_t)in.u32()});
      }
    gatts_ = in.atts();
    tag = in.u32();
    n = in.u32();
    if (tag == kTagVar)
      for (uint32_t k = 0; k < n && !in.bad; ++k) {
        VarInfo v;
        v.name = in.name();
        const uint32_t nd = in.u32();
        v.count = 1;
        for (uint32_t d = 0; d < nd && !in.bad; ++d) {
          const int id = (int)in.u32();
          if (id < 0 || id >= (int)dims_.size()) {
            in.bad = true;
            break;
          }
          v.dims.push_back(id);
          v.count *= dims_[id].second;
        }
        v.atts = in.atts();
        v.type = (Type)in.u32();
        in.u32();  // vsize
        v.begin = offset64 ? (int64_t)in.u64() : (int64_t)in.u32();
        vars_[v.name] = v;
      }
    if (!in.bad) {
      // every variable's data must lie inside the file
      f.clear();
      f.seekg(0, std::ios::end);
      const int64_t size = (int64_t)f.tellg();
      for (const auto& kv : vars_) {
        const VarInfo& v = kv.second;
        if (v.begin < 0 || v.count < 0 || v.begin + v.count * type_size(v.type) > size) {
          err = path + ": truncated (variable " + v.name + " ends past the end of the file)";
          return false;
        }
      }
      return true;
    }
    if ((size_t)f.gcount() < want) {
      err = path + ": truncated or malformed header";
      return false;
    }
  }
}

int64_t Reader::dim(const std::string& name) const {
  for (const auto& d : dims_)
    if (d.first == name) return d.second;
  return -1;
}

const Att* Reader::gatt(const std::string& name) const {
  for (const Att& a : gatts_)
    if (a.name == name) return &a;
  return nullptr;
}

bool Reader::raw(const VarInfo& v, std::vector<uint8_t>& bytes) const {
  std::ifstream f(path_, std::ios::binary);
  if (!f) return false;
  const int sz = type_size(v.type);
  bytes.assign((size_t)(v.count * sz), 0);
  f.seekg(v.begin);
  f.read((char*)bytes.data(), (std::streamsize)bytes.size());
  if ((int64_t)f.gcount() != v.count * sz) return false;
  for (int64_t i = 0; i < v.count; ++i)  // from big-endian
    for (int c = 0; c < sz / 2; ++c) std::swap(bytes[(size_t)(i * sz + c)], bytes[(size_t)(i * sz + sz - 1 - c)]);
  return true;
}

bool Reader::read(const std::string& name, std::vector<double>& out) const {
  auto it = vars_.find(name);
  if (it == vars_.end()) return false;
  std::vector<uint8_t> b;
  if (!raw(it->second, b)) return false;
  const VarInfo& v = it->second;
  out.resize((size_t)v.count);
  for (int64_t i = 0; i < v.count; ++i) {
    if (v.type == kDouble) {
      std::memcpy(&out[i], &b[(size_t)(8 * i)], 8);
    } else if (v.type == kInt) {
      int32_t x;
      std::memcpy(&x, &b[(size_t)(4 * i)], 4);
      out[i] = x;
    } else {
      out[i] = (int8_t)b[(size_t)i];
    }
  }
  return true;
}

bool Reader::read(const std::string& name, std::vector<int32_t>& out) const {
  auto it = vars_.find(name);
  if (it == vars_.end() || it->second.type == kDouble) return false;
  std::vector<uint8_t> b;
  if (!raw(it->second, b)) return false;
  const VarInfo& v = it->second;
  out.resize((size_t)v.count);
  for (int64_t i = 0; i < v.count; ++i) {
    if (v.type == kInt)
      std::memcpy(&out[i], &b[(size_t)(4 * i)], 4);
    else
      out[i] = (int8_t)b[(size_t)i];
  }
  return true;
}

bool Reader::read(const std::string& name, std::vector<uint8_t>& out) const {
  auto it = vars_.find(name);
  if (it == vars_.end() || (it->second.type != kByte && it->second.type != kChar)) return false;
  return raw(it->second, out);
}

}  // namespace nc
}  // namespace lpm
