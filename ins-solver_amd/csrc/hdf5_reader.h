// Minimal reader of the HDF5 files OP2 writes for its meshes (op_decl_set_hdf5 /
// op_decl_map_hdf5 / op_decl_dat_hdf5): superblock version 0 or 1 with 8-byte addresses,
// version-1 object headers (with continuation blocks), the root group as a symbol table
// (version-1 B-tree of symbol-table nodes plus a local heap), and datasets of 32-bit
// integers or 64-bit floats, little-endian, stored contiguously or compactly without
// filters. Anything else is reported as unsupported. No HDF5 library is involved.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace insp {

class Hdf5File {
 public:
  explicit Hdf5File(const std::string& path);
  bool has(const std::string& name) const { return root_.count(name) != 0; }
  std::vector<std::string> names() const;
  // Dataset `name` as 32-bit integers; dims receives its shape.
  std::vector<int32_t> read_int(const std::string& name, std::vector<uint64_t>* dims = nullptr) const;
  std::vector<double> read_double(const std::string& name, std::vector<uint64_t>* dims = nullptr) const;

 private:
  struct Dataset {
    std::vector<uint64_t> dims;
    int cls = -1;       // 0 fixed point, 1 floating point
    int size = 0;       // bytes per element
    bool is_signed = true;
    uint64_t addr = 0;  // contiguous data address (or offset of compact data in the file)
    uint64_t bytes = 0;
  };
  const uint8_t* at(uint64_t off, uint64_t len) const;
  uint64_t u64(uint64_t off) const;
  uint32_t u32(uint64_t off) const;
  uint16_t u16(uint64_t off) const;
  uint8_t u8(uint64_t off) const;
  std::string cstr(uint64_t off) const;
  struct Message {
    uint16_t type;
    uint16_t size;
    uint64_t body;
  };
  std::vector<Message> messages(uint64_t ohdr) const;
  void read_group(uint64_t btree, uint64_t heap);
  Dataset dataset(const std::string& name) const;

  std::string path_;
  std::vector<uint8_t> d_;
  std::map<std::string, uint64_t> root_;  // name -> object header address
};

}  // namespace insp
