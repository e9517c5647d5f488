#include "hdf5_reader.h"

#include <cstring>
#include <fstream>
#include <stdexcept>

namespace insp {

namespace {
const uint8_t kSignature[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
}

Hdf5File::Hdf5File(const std::string& path) : path_(path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::invalid_argument("cannot open " + path);
  const std::streamsize n = f.tellg();
  if (n < 96) throw std::invalid_argument(path + " is not an HDF5 file");
  d_.resize((size_t)n);
  f.seekg(0);
  if (!f.read(reinterpret_cast<char*>(d_.data()), n)) throw std::invalid_argument("cannot read " + path);
  // the superblock sits at 0 or after a user block of 512, 1024, 2048, ... bytes
  uint64_t sb = UINT64_MAX;
  for (uint64_t off = 0; off + 8 <= d_.size(); off = off ? off * 2 : 512)
    if (std::memcmp(&d_[off], kSignature, 8) == 0) { sb = off; break; }
  if (sb == UINT64_MAX) throw std::invalid_argument(path + " is not an HDF5 file");
  const int version = u8(sb + 8);
  if (version > 1)
    throw std::invalid_argument(path + ": HDF5 superblock version " + std::to_string(version) +
                                " is not supported (versions 0 and 1 are)");
  if (u8(sb + 13) != 8 || u8(sb + 14) != 8)
    throw std::invalid_argument(path + ": only 8-byte addresses and lengths are supported");
  const uint64_t base_off = sb + (version == 0 ? 24 : 28);
  if (u64(base_off) != 0 && u64(base_off) != sb)
    throw std::invalid_argument(path + ": unexpected HDF5 base address");
  const uint64_t root_entry = base_off + 32;
  const uint64_t root_ohdr = u64(root_entry + 8);
  uint64_t btree = 0, heap = 0;
  for (const Message& m : messages(root_ohdr))
    if (m.type == 0x11) {  // symbol table message
      btree = u64(m.body);
      heap = u64(m.body + 8);
    }
  if (!btree || !heap)
    throw std::invalid_argument(path + ": the root group is not a symbol-table group");
  read_group(btree, heap);
}

const uint8_t* Hdf5File::at(uint64_t off, uint64_t len) const {
  if (off > d_.size() || len > d_.size() - off)
    throw std::invalid_argument(path_ + ": HDF5 structure points outside the file");
  return d_.data() + off;
}
uint64_t Hdf5File::u64(uint64_t off) const { uint64_t v; std::memcpy(&v, at(off, 8), 8); return v; }
uint32_t Hdf5File::u32(uint64_t off) const { uint32_t v; std::memcpy(&v, at(off, 4), 4); return v; }
uint16_t Hdf5File::u16(uint64_t off) const { uint16_t v; std::memcpy(&v, at(off, 2), 2); return v; }
uint8_t Hdf5File::u8(uint64_t off) const { return *at(off, 1); }

std::string Hdf5File::cstr(uint64_t off) const {
  std::string s;
  for (uint64_t p = off;; ++p) {
    const char c = (char)u8(p);
    if (!c) break;
    s.push_back(c);
    if (s.size() > 4096) throw std::invalid_argument(path_ + ": unterminated name");
  }
  return s;
}

// Messages of a version-1 object header, following continuation blocks.
std::vector<Hdf5File::Message> Hdf5File::messages(uint64_t ohdr) const {
  if (u8(ohdr) != 1)
    throw std::invalid_argument(path_ + ": object header version " + std::to_string(u8(ohdr)) +
                                " is not supported (version 1 is)");
  const int nmsg = u16(ohdr + 2);
  std::vector<std::pair<uint64_t, uint64_t>> blocks{{ohdr + 16, u32(ohdr + 8)}};
  std::vector<Message> out;
  for (size_t b = 0; b < blocks.size() && (int)out.size() < nmsg; ++b) {
    if (blocks.size() > 64) throw std::invalid_argument(path_ + ": too many header continuations");
    uint64_t p = blocks[b].first;
    const uint64_t end = blocks[b].first + blocks[b].second;
    while (p + 8 <= end && (int)out.size() < nmsg) {
      Message m{u16(p), u16(p + 2), p + 8};
      at(m.body, m.size);
      if (m.type == 0x10) blocks.push_back({u64(m.body), u64(m.body + 8)});
      out.push_back(m);
      p = m.body + m.size;
    }
  }
  return out;
}

void Hdf5File::read_group(uint64_t btree, uint64_t heap) {
  if (std::memcmp(at(heap, 4), "HEAP", 4) != 0) throw std::invalid_argument(path_ + ": bad local heap");
  const uint64_t heap_data = u64(heap + 24);
  std::vector<uint64_t> stack{btree};
  int visited = 0;
  while (!stack.empty()) {
    const uint64_t node = stack.back();
    stack.pop_back();
    if (++visited > 100000) throw std::invalid_argument(path_ + ": group B-tree too large");
    if (std::memcmp(at(node, 4), "TREE", 4) != 0) throw std::invalid_argument(path_ + ": bad group B-tree");
    if (u8(node + 4) != 0) throw std::invalid_argument(path_ + ": not a group B-tree");
    const int level = u8(node + 5), used = u16(node + 6);
    uint64_t p = node + 24;  // signature, type, level, entries used, two sibling addresses
    for (int i = 0; i < used; ++i) {
      p += 8;  // key: heap offset
      const uint64_t child = u64(p);
      p += 8;
      if (level > 0) {
        stack.push_back(child);
        continue;
      }
      if (std::memcmp(at(child, 4), "SNOD", 4) != 0) throw std::invalid_argument(path_ + ": bad symbol node");
      const int n = u16(child + 6);
      for (int j = 0; j < n; ++j) {
        const uint64_t e = child + 8 + 40 * (uint64_t)j;
        root_[cstr(heap_data + u64(e))] = u64(e + 8);
      }
    }
  }
}

std::vector<std::string> Hdf5File::names() const {
  std::vector<std::string> v;
  for (auto& kv : root_) v.push_back(kv.first);
  return v;
}

Hdf5File::Dataset Hdf5File::dataset(const std::string& name) const {
  auto it = root_.find(name);
  if (it == root_.end()) throw std::invalid_argument(path_ + ": no dataset " + name);
  Dataset ds;
  bool have_space = false, have_layout = false;
  for (const Message& m : messages(it->second)) {
    if (m.type == 0x1) {  // dataspace
      const int ver = u8(m.body), rank = u8(m.body + 1);
      if (ver != 1 && ver != 2) throw std::invalid_argument(path_ + ": unsupported dataspace of " + name);
      const uint64_t dims = m.body + (ver == 1 ? 8 : 4);
      for (int r = 0; r < rank; ++r) ds.dims.push_back(u64(dims + 8 * (uint64_t)r));
      have_space = true;
    } else if (m.type == 0x3) {  // datatype
      const uint8_t cv = u8(m.body), bits = u8(m.body + 1);
      ds.cls = cv & 15;
      ds.size = (int)u32(m.body + 4);
      if (bits & 1) throw std::invalid_argument(path_ + ": big-endian data in " + name);
      ds.is_signed = (bits & 8) != 0;
    } else if (m.type == 0x8) {  // data layout
      const int ver = u8(m.body);
      if (ver == 3) {
        const int cls = u8(m.body + 1);
        if (cls == 0) {
          ds.bytes = u16(m.body + 2);
          ds.addr = m.body + 4;
        } else if (cls == 1) {
          ds.addr = u64(m.body + 2);
          ds.bytes = u64(m.body + 10);
        } else {
          throw std::invalid_argument(path_ + ": chunked storage of " + name + " is not supported");
        }
      } else if (ver == 1 || ver == 2) {
        const int rank = u8(m.body + 1), cls = u8(m.body + 2);
        if (cls == 1) {
          ds.addr = u64(m.body + 8);
          ds.bytes = UINT64_MAX;  // from the dataspace
        } else if (cls == 0) {
          const uint64_t sz = m.body + 8 + 4 * (uint64_t)rank;
          ds.bytes = u32(sz);
          ds.addr = sz + 4;
        } else {
          throw std::invalid_argument(path_ + ": chunked storage of " + name + " is not supported");
        }
      } else {
        throw std::invalid_argument(path_ + ": unsupported layout of " + name);
      }
      have_layout = true;
    } else if (m.type == 0xB) {
      throw std::invalid_argument(path_ + ": filtered (compressed) data in " + name + " is not supported");
    }
  }
  if (!have_space || !have_layout || ds.cls < 0) throw std::invalid_argument(path_ + ": incomplete dataset " + name);
  uint64_t count = 1;
  for (uint64_t v : ds.dims) {
    if (v && count > UINT64_MAX / v) throw std::invalid_argument(path_ + ": dataset too large");
    count *= v;
  }
  const uint64_t need = count * (uint64_t)ds.size;
  if (ds.bytes == UINT64_MAX) ds.bytes = need;
  if (ds.bytes < need) throw std::invalid_argument(path_ + ": dataset " + name + " is truncated");
  at(ds.addr, need);
  return ds;
}

std::vector<int32_t> Hdf5File::read_int(const std::string& name, std::vector<uint64_t>* dims) const {
  const Dataset ds = dataset(name);
  if (ds.cls != 0 || ds.size != 4) throw std::invalid_argument(path_ + ": " + name + " is not 32-bit integer data");
  uint64_t count = 1;
  for (uint64_t v : ds.dims) count *= v;
  std::vector<int32_t> out((size_t)count);
  std::memcpy(out.data(), at(ds.addr, count * 4), count * 4);
  if (dims) *dims = ds.dims;
  return out;
}

std::vector<double> Hdf5File::read_double(const std::string& name, std::vector<uint64_t>* dims) const {
  const Dataset ds = dataset(name);
  if (ds.cls != 1 || ds.size != 8) throw std::invalid_argument(path_ + ": " + name + " is not 64-bit float data");
  uint64_t count = 1;
  for (uint64_t v : ds.dims) count *= v;
  std::vector<double> out((size_t)count);
  std::memcpy(out.data(), at(ds.addr, count * 8), count * 8);
  if (dims) *dims = ds.dims;
  return out;
}

}  // namespace insp
