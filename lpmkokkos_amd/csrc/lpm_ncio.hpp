// NetCDF classic files (the 64-bit-offset variant, "CDF\x02"), written and
// read without the netCDF library, which this image does not have.  Only what
// the mesh and BVE checkpoints need: fixed-size dimensions, global and
// per-variable attributes (text, int, double), and variables of type byte,
// int and double.  Files follow the published classic format specification
// (big-endian header and data, 4-byte alignment) and open in any netCDF
// reader; the variable layout of a mesh file is that of
// NcWriter::writePolymesh (src/LpmNetCDF_Impl.hpp:46-218).
#pragma once

#include <cstdint>
#include <list>
#include <map>
#include <string>
#include <vector>

namespace lpm {
namespace nc {

enum Type : int32_t { kByte = 1, kChar = 2, kInt = 4, kDouble = 6 };

struct Att {
  std::string name;
  Type type = kChar;
  std::string text;            // kChar
  std::vector<int32_t> ints;   // kInt
  std::vector<double> doubles; // kDouble
};

struct VarInfo {
  std::string name;
  Type type = kDouble;
  std::vector<int> dims;  // dimension ids
  std::vector<Att> atts;
  int64_t begin = 0;      // file offset of the data
  int64_t count = 0;      // elements
};

class Writer {
 public:
  int dim(const std::string& name, int64_t len);  // returns the dimension id
  void att_text(const std::string& name, const std::string& v);
  void att_int(const std::string& name, int32_t v);
  void att_double(const std::string& name, double v);
  // data (host byte order) must stay valid until write(); count = product of
  // the dimensions; a non-null `units` adds a "units" text attribute
  void var(const std::string& name, Type type, const std::vector<int>& dims, const void* data,
           const char* units = nullptr);
  // as var(), with the data copied into the writer
  void var_copy(const std::string& name, Type type, const std::vector<int>& dims, const void* data,
                const char* units = nullptr);
  // false with `err` set on failure
  bool write(const std::string& path, std::string& err) const;

 private:
  std::vector<std::pair<std::string, int64_t>> dims_;
  std::vector<Att> gatts_;
  std::vector<VarInfo> vars_;
  std::vector<const void*> data_;
  std::list<std::vector<uint8_t>> owned_;
};

class Reader {
 public:
  bool open(const std::string& path, std::string& err);
  bool has_var(const std::string& name) const { return vars_.count(name) != 0; }
  int64_t dim(const std::string& name) const;  // -1 when absent
  const Att* gatt(const std::string& name) const;
  // whole variable, converted from its stored type; false if absent
  bool read(const std::string& name, std::vector<double>& out) const;
  bool read(const std::string& name, std::vector<int32_t>& out) const;
  bool read(const std::string& name, std::vector<uint8_t>& out) const;

 private:
  std::string path_;
  std::vector<std::pair<std::string, int64_t>> dims_;
  std::vector<Att> gatts_;
  std::map<std::string, VarInfo> vars_;
  bool raw(const VarInfo& v, std::vector<uint8_t>& bytes) const;
};

}  // namespace nc
}  // namespace lpm
