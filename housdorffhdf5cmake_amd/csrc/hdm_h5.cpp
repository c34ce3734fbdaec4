// HDF5 volume loader (include/hdm_h5.h): the library's own reader of the
// HDF5 file format, for the plain numeric datasets that the reference's
// benchmark loads (forBench/hdf5Manag.cu:44-83, 196-220).
//
// Structure walk: superblock -> root group object header -> (symbol table
// B-tree + local heap, or link messages) per path component -> the
// dataset's object header, whose dataspace, datatype, layout and filter
// pipeline messages say where the elements are and how they are coded.
// Field layouts follow the HDF5 File Format Specification (version 3.0):
// superblock (II.A), object headers (IV.A.1), messages (IV.A.2), B-tree v1
// (III.A.1), symbol table nodes (III.B), local heaps (III.D), fixed-array
// chunk index (III.H... as written by HDF5 1.10 for layout version 4).
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hdm_h5.h"

namespace {

struct H5Err {
  int code;
};
[[noreturn]] void fail(int code) { throw H5Err{code}; }
[[noreturn]] void bad() { fail(HDM_E_FORMAT); }
[[noreturn]] void unsupported() { fail(HDM_E_UNSUPPORTED); }

struct File {
  int fd = -1;
  mutable uint64_t nodes = 0;  // B-tree / heap nodes visited (bounds corrupt files)
  void visit() const {
    if (++nodes > (1u << 22)) bad();
  }
  uint64_t size = 0;
  uint64_t base = 0;  // absolute address that file addresses are relative to
  int so = 8, sl = 8;  // bytes of an offset (address) and of a length
  uint64_t root = 0;   // root group object header
  ~File() {
    if (fd >= 0) close(fd);
  }
  uint64_t undef() const { return so == 8 ? ~0ull : ((1ull << (8 * so)) - 1); }
  void read_abs(uint64_t off, void *dst, uint64_t n) const {
    if (off > size || n > size - off) bad();
    uint8_t *d = (uint8_t *)dst;
    while (n) {
      const ssize_t r = pread(fd, d, n > (1u << 30) ? (1u << 30) : n, (off_t)off);
      if (r <= 0) fail(HDM_E_IO);
      d += r, off += (uint64_t)r, n -= (uint64_t)r;
    }
  }
  void read(uint64_t addr, void *dst, uint64_t n) const { read_abs(base + addr, dst, n); }
  std::vector<uint8_t> bytes(uint64_t addr, uint64_t n) const {
    // sizes come from the file: check them against it before allocating
    if (base + addr > size || n > size - (base + addr)) bad();
    std::vector<uint8_t> b(n);
    read(addr, b.data(), n);
    return b;
  }
};

// little-endian field reader over a byte range
struct Cur {
  const uint8_t *p, *e;
  Cur(const uint8_t *b, size_t n) : p(b), e(b + n) {}
  explicit Cur(const std::vector<uint8_t> &v) : p(v.data()), e(v.data() + v.size()) {}
  uint64_t u(int n) {
    if (n < 0 || n > 8 || (size_t)(e - p) < (size_t)n) bad();
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v |= (uint64_t)p[i] << (8 * i);
    p += n;
    return v;
  }
  uint8_t u8() { return (uint8_t)u(1); }
  void skip(uint64_t n) {
    if ((uint64_t)(e - p) < n) bad();
    p += n;
  }
  size_t left() const { return (size_t)(e - p); }
};

struct Msg {
  uint16_t type;
  uint8_t flags;
  std::vector<uint8_t> data;
};

enum : uint16_t {
  kDataspace = 1,
  kLinkInfo = 2,
  kDatatype = 3,
  kLink = 6,
  kLayout = 8,
  kFilters = 11,
  kContinuation = 16,
  kSymbolTable = 17,
};

void sig(const File &f, uint64_t addr, const char *s) {
  char b[4];
  f.read(addr, b, 4);
  if (std::memcmp(b, s, 4) != 0) bad();
}

// Messages of the object header at addr, continuation blocks followed
// (IV.A.1.a version 1, IV.A.1.b version 2).
std::vector<Msg> object_header(const File &f, uint64_t addr) {
  std::vector<Msg> out;
  struct Block {
    uint64_t addr, len;
  };
  std::vector<Block> blocks;
  uint8_t h[16];
  f.read(addr, h, 4);
  int version;
  int v2_flags = 0;
  uint64_t nmsgs = ~0ull;
  if (std::memcmp(h, "OHDR", 4) == 0) {
    version = 2;
    uint8_t b[2];
    f.read(addr + 4, b, 2);
    if (b[0] != 2) unsupported();
    v2_flags = b[1];
    uint64_t o = addr + 6;
    if (v2_flags & 0x20) o += 16;  // times
    if (v2_flags & 0x10) o += 4;   // attribute phase change values
    const int szb = 1 << (v2_flags & 3);
    uint8_t sz[8] = {0};
    f.read(o, sz, szb);
    uint64_t chunk0 = 0;
    for (int i = 0; i < szb; ++i) chunk0 |= (uint64_t)sz[i] << (8 * i);
    o += szb;
    blocks.push_back({o, chunk0});
  } else {
    f.read(addr, h, 16);
    version = h[0];
    if (version != 1) unsupported();
    nmsgs = (uint64_t)h[2] | ((uint64_t)h[3] << 8);
    const uint64_t hsize = (uint64_t)h[8] | ((uint64_t)h[9] << 8) | ((uint64_t)h[10] << 16) |
                           ((uint64_t)h[11] << 24);
    blocks.push_back({addr + 16, hsize});
  }
  for (size_t bi = 0; bi < blocks.size() && bi < 4096; ++bi) {
    Block blk = blocks[bi];
    if (version == 2 && bi > 0) {  // continuation chunk: "OCHK" ... checksum
      sig(f, blk.addr, "OCHK");
      if (blk.len < 8) bad();
      blk.addr += 4, blk.len -= 8;
    }
    const std::vector<uint8_t> raw = f.bytes(blk.addr, blk.len);
    Cur c(raw);
    const size_t hdr = version == 1 ? 8 : (4 + ((v2_flags & 0x04) ? 2 : 0));
    while (c.left() >= hdr && out.size() < nmsgs) {
      Msg m;
      uint64_t size;
      if (version == 1) {
        m.type = (uint16_t)c.u(2);
        size = c.u(2);
        m.flags = c.u8();
        c.skip(3);
      } else {
        m.type = c.u8();
        size = c.u(2);
        m.flags = c.u8();
        if (v2_flags & 0x04) c.skip(2);  // creation order
      }
      if (c.left() < size) bad();
      m.data.assign(c.p, c.p + size);
      c.skip(size);
      if (m.type == kContinuation) {
        Cur d(m.data);
        const uint64_t a = d.u(f.so), l = d.u(f.sl);
        blocks.push_back({a, l});
      }
      out.push_back(std::move(m));
    }
  }
  return out;
}

const Msg *find_msg(const std::vector<Msg> &ms, uint16_t type) {
  for (const Msg &m : ms)
    if (m.type == type) {
      if (m.flags & 0x02) unsupported();  // shared message (committed type, SOHM)
      return &m;
    }
  return nullptr;
}

// Child `name` of a group stored as a symbol table: a B-tree (type 0) whose
// leaves are symbol table nodes, names in the local heap (III.A.1, III.B,
// III.D).
uint64_t symtab_lookup(const File &f, uint64_t btree, uint64_t heap, const std::string &name,
                       int depth = 0, int want_level = -1) {
  if (depth > 64) bad();
  f.visit();
  // local heap: data segment address
  sig(f, heap, "HEAP");
  std::vector<uint8_t> hh = f.bytes(heap + 8, f.sl * 2 + f.so);
  Cur hc(hh);
  const uint64_t seg_size = hc.u(f.sl);
  hc.u(f.sl);
  const uint64_t seg = hc.u(f.so);
  const std::vector<uint8_t> names = f.bytes(seg, seg_size);
  auto name_at = [&](uint64_t off) -> std::string {
    if (off >= names.size()) bad();
    const char *s = (const char *)names.data() + off;
    return std::string(s, strnlen(s, names.size() - off));
  };
  // B-tree walk
  sig(f, btree, "TREE");
  std::vector<uint8_t> nh = f.bytes(btree + 4, 4 + 2 * f.so);
  Cur nc(nh);
  if (nc.u8() != 0) bad();  // group node
  const int level = nc.u8();
  if (want_level >= 0 && level != want_level) bad();
  const uint64_t entries = nc.u(2);
  const uint64_t kv = 4 + 4 + 2 * f.so;
  const std::vector<uint8_t> body = f.bytes(btree + kv, (entries + 1) * f.sl + entries * f.so);
  Cur bc(body);
  for (uint64_t i = 0; i < entries; ++i) {
    bc.u(f.sl);  // key: heap offset of the greatest name in the left subtree
    const uint64_t child = bc.u(f.so);
    if (level > 0) {
      const uint64_t r = symtab_lookup(f, child, heap, name, depth + 1, level - 1);
      if (r != ~0ull) return r;
      continue;
    }
    sig(f, child, "SNOD");
    uint8_t sh[4];
    f.read(child + 4, sh, 4);
    const uint64_t nsym = (uint64_t)sh[2] | ((uint64_t)sh[3] << 8);
    const uint64_t esz = 2 * (uint64_t)f.so + 24;
    const std::vector<uint8_t> ents = f.bytes(child + 8, nsym * esz);
    Cur ec(ents);
    for (uint64_t k = 0; k < nsym; ++k) {
      const uint64_t noff = ec.u(f.so);
      const uint64_t ohdr = ec.u(f.so);
      ec.skip(24);
      if (name_at(noff) == name) return ohdr;
    }
  }
  return ~0ull;
}

// Child `name` of the group whose object header messages are ms.
uint64_t group_child(const File &f, const std::vector<Msg> &ms, const std::string &name) {
  for (const Msg &m : ms) {
    if (m.type == kSymbolTable) {
      Cur c(m.data);
      const uint64_t bt = c.u(f.so), heap = c.u(f.so);
      return symtab_lookup(f, bt, heap, name);
    }
  }
  for (const Msg &m : ms) {
    if (m.type == kLinkInfo) {
      Cur c(m.data);
      c.u8();
      const int flags = c.u8();
      if (flags & 1) c.skip(8);
      const uint64_t heap = c.u(f.so);
      if (heap != f.undef()) unsupported();  // dense link storage (fractal heap)
    }
  }
  for (const Msg &m : ms) {
    if (m.type != kLink) continue;
    Cur c(m.data);
    if (c.u8() != 1) unsupported();
    const int flags = c.u8();
    int ltype = 0;
    if (flags & 0x08) ltype = c.u8();
    if (flags & 0x04) c.skip(8);
    if (flags & 0x10) c.skip(1);
    const uint64_t nlen = c.u(1 << (flags & 3));
    if (c.left() < nlen) bad();
    const std::string ln((const char *)c.p, nlen);
    c.skip(nlen);
    if (ln != name) continue;
    if (ltype != 0) unsupported();  // soft or external link
    return c.u(f.so);
  }
  return ~0ull;
}

void open_file(File &f, const char *path) {
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) fail(HDM_E_IO);
  struct stat st;
  if (fstat(f.fd, &st) != 0) fail(HDM_E_IO);
  f.size = (uint64_t)st.st_size;
  static const uint8_t kSig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  uint64_t sb = ~0ull;
  for (uint64_t off = 0; off + 64 <= f.size; off = off ? off * 2 : 512) {
    uint8_t b[8];
    f.read_abs(off, b, 8);
    if (std::memcmp(b, kSig, 8) == 0) {
      sb = off;
      break;
    }
  }
  if (sb == ~0ull) bad();
  uint8_t h[256];
  const uint64_t n = std::min<uint64_t>(sizeof(h), f.size - sb);
  f.read_abs(sb, h, n);
  Cur c(h + 8, n - 8);
  const int ver = c.u8();
  if (ver <= 1) {
    c.skip(4);  // free space, root group, reserved, shared header versions
    f.so = c.u8();
    f.sl = c.u8();
    c.skip(1);
    c.skip(4);  // group leaf / internal node K
    c.skip(4);  // consistency flags
    if (ver == 1) c.skip(4);
    if (f.so < 2 || f.so > 8 || f.sl < 2 || f.sl > 8) bad();
    f.base = c.u(f.so);
    c.u(f.so);  // free space info
    c.u(f.so);  // end of file
    c.u(f.so);  // driver info
    c.u(f.so);  // root: link name offset
    f.root = c.u(f.so);
  } else if (ver == 2 || ver == 3) {
    f.so = c.u8();
    f.sl = c.u8();
    c.skip(1);
    if (f.so < 2 || f.so > 8 || f.sl < 2 || f.sl > 8) bad();
    f.base = c.u(f.so);
    c.u(f.so);  // superblock extension
    c.u(f.so);  // end of file
    f.root = c.u(f.so);
  } else {
    unsupported();
  }
}

struct DType {
  int cls = -1;  // HDM_H5_*
  int size = 0;
  bool be = false, sign = false;
};

DType parse_dtype(Cur &c, int depth = 0) {
  if (depth > 4) bad();
  const uint8_t b0 = c.u8();
  const int cls = b0 & 0x0F;
  const uint32_t bits = (uint32_t)c.u(3);
  DType t;
  t.size = (int)c.u(4);
  t.be = bits & 1;
  switch (cls) {
    case 0: {  // fixed-point
      const uint64_t off = c.u(2), prec = c.u(2);
      if (off != 0 || prec != 8ull * t.size) unsupported();
      t.cls = (bits & 8) ? HDM_H5_INT : HDM_H5_UINT;
      t.sign = bits & 8;
      break;
    }
    case 1: {  // floating point
      if ((bits & 0x41) == 0x41 || (bits & 0x40)) unsupported();  // VAX order
      c.skip(12);
      if (t.size != 4 && t.size != 8) unsupported();
      t.cls = HDM_H5_FLOAT;
      break;
    }
    case 4: {  // bitfield
      const uint64_t off = c.u(2), prec = c.u(2);
      if (off != 0 || prec != 8ull * t.size) unsupported();
      t.cls = HDM_H5_UINT;
      break;
    }
    case 8: {  // enumeration: the base type follows (names and values unused)
      const DType base = parse_dtype(c, depth + 1);
      if (base.cls != HDM_H5_INT && base.cls != HDM_H5_UINT) unsupported();
      t = base;
      t.cls = HDM_H5_ENUM;
      break;
    }
    default:
      unsupported();
  }
  if (t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8) unsupported();
  return t;
}

struct Filter {
  int id;
  std::vector<uint32_t> cd;
};

struct Dataset {
  int rank = 0;
  uint64_t dims[8] = {0};
  uint64_t n = 1;
  DType t;
  int layout = -1;
  int layout_version = 0;
  uint64_t addr = ~0ull, size = 0;  // contiguous data, or the chunk index
  std::vector<uint8_t> compact;
  uint64_t chunk[8] = {0};
  int index = 0;  // chunked: 0 B-tree v1, 1 single, 2 implicit, 3 fixed array
  uint64_t single_size = 0;
  uint32_t single_mask = 0;
  bool single_filtered = false;
  std::vector<Filter> filters;
};

std::vector<std::string> split_path(const char *p) {
  std::vector<std::string> out;
  std::string cur;
  for (const char *s = p; *s; ++s) {
    if (*s == '/') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += *s;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

Dataset open_dataset(File &f, const char *path, const char *name) {
  if (!path || !name) fail(HDM_E_INVALID);
  open_file(f, path);
  uint64_t oh = f.root;
  const std::vector<std::string> parts = split_path(name);
  if (parts.empty()) fail(HDM_E_NOTFOUND);
  std::vector<Msg> ms = object_header(f, oh);
  for (const std::string &pc : parts) {
    oh = group_child(f, ms, pc);
    if (oh == ~0ull) fail(HDM_E_NOTFOUND);
    ms = object_header(f, oh);
  }
  Dataset d;
  const Msg *ds = find_msg(ms, kDataspace);
  const Msg *dt = find_msg(ms, kDatatype);
  const Msg *ly = find_msg(ms, kLayout);
  if (!ds || !dt || !ly) fail(HDM_E_NOTFOUND);  // a group, not a dataset
  {  // dataspace (IV.A.2.b)
    Cur c(ds->data);
    const int ver = c.u8();
    d.rank = c.u8();
    c.u8();  // flags
    if (ver == 1) {
      c.skip(5);
    } else if (ver == 2) {
      const int type = c.u8();
      if (type == 2) d.rank = 0, d.n = 0;  // null dataspace
    } else {
      unsupported();
    }
    if (d.rank > 8) unsupported();
    for (int i = 0; i < d.rank; ++i) {
      d.dims[i] = c.u(f.sl);
      if (__builtin_mul_overflow(d.n, d.dims[i], &d.n) || d.n > (1ull << 48)) bad();
    }
  }
  {  // datatype (IV.A.2.d)
    Cur c(dt->data);
    d.t = parse_dtype(c);
  }
  if (const Msg *fp = find_msg(ms, kFilters)) {  // filter pipeline (IV.A.2.l)
    Cur c(fp->data);
    const int ver = c.u8();
    const int nf = c.u8();
    if (ver == 1) c.skip(6);
    else if (ver != 2) unsupported();
    for (int i = 0; i < nf; ++i) {
      Filter flt;
      flt.id = (int)c.u(2);
      uint64_t nlen = 0;
      if (ver == 1 || flt.id >= 256) nlen = c.u(2);
      c.u(2);  // flags
      const uint64_t ncd = c.u(2);
      if (nlen) c.skip(ver == 1 ? (nlen + 7) / 8 * 8 : nlen);
      for (uint64_t k = 0; k < ncd; ++k) flt.cd.push_back((uint32_t)c.u(4));
      if (ver == 1 && (ncd & 1)) c.skip(4);
      if (flt.id != 1 && flt.id != 2 && flt.id != 3) unsupported();
      d.filters.push_back(flt);
    }
  }
  {  // data layout (IV.A.2.i)
    Cur c(ly->data);
    const int ver = c.u8();
    d.layout_version = ver;
    if (ver == 3 || ver == 4) {
      d.layout = c.u8();
      if (d.layout == 0) {
        const uint64_t sz = c.u(2);
        if (c.left() < sz) bad();
        d.compact.assign(c.p, c.p + sz);
      } else if (d.layout == 1) {
        d.addr = c.u(f.so);
        d.size = c.u(f.sl);
      } else if (d.layout == 2) {
        if (ver == 3) {
          const int nd = c.u8();
          if (nd != d.rank + 1) bad();
          d.addr = c.u(f.so);
          for (int i = 0; i < nd; ++i) {
            const uint64_t v = c.u(4);
            if (i < d.rank) d.chunk[i] = v;
          }
          d.index = 0;
        } else {
          const int flags = c.u8();
          const int nd = c.u8();
          const int enc = c.u8();
          if (nd != d.rank + 1 || enc < 1 || enc > 8) bad();
          for (int i = 0; i < nd; ++i) {
            const uint64_t v = c.u(enc);
            if (i < d.rank) d.chunk[i] = v;
          }
          const int it = c.u8();
          if (it == 1) {
            d.index = 1;
            if (flags & 2) {
              d.single_filtered = true;
              d.single_size = c.u(f.sl);
              d.single_mask = (uint32_t)c.u(4);
            }
          } else if (it == 2) {
            d.index = 2;
          } else if (it == 3) {
            d.index = 3;
            c.u8();  // page bits (repeated in the index header)
          } else {
            unsupported();  // extensible array, B-tree v2 (unlimited dimensions)
          }
          d.addr = c.u(f.so);
        }
      } else {
        unsupported();  // virtual
      }
    } else if (ver == 1 || ver == 2) {
      const int nd = c.u8();
      d.layout = c.u8();
      c.skip(5);
      if (d.layout != 0) d.addr = c.u(f.so);
      std::vector<uint64_t> dd;
      for (int i = 0; i < nd; ++i) dd.push_back(c.u(4));
      if (d.layout == 2) {
        if (nd != d.rank + 1) bad();
        for (int i = 0; i < d.rank; ++i) d.chunk[i] = dd[i];
        d.index = 0;
      } else if (d.layout == 1) {
        d.size = d.n * (uint64_t)d.t.size;
      } else {
        const uint64_t sz = c.u(4);
        if (c.left() < sz) bad();
        d.compact.assign(c.p, c.p + sz);
      }
    } else {
      unsupported();
    }
  }
  if (d.layout == 2)
    for (int i = 0; i < d.rank; ++i)
      if (d.chunk[i] == 0) bad();
  return d;
}

// ---------------------------------------------------------------- decode
// element k of a coded run -> the destination type
struct Conv {
  DType t;
  int mode;
  template <typename T>
  static T bswap(T v) {
    uint8_t *b = (uint8_t *)&v;
    std::reverse(b, b + sizeof(T));
    return v;
  }
  // converts n elements from src into dst (element index base of dst);
  // returns false when a value does not fit
  bool run(const uint8_t *src, uint64_t n, void *dst, uint64_t at) const {
    const int s = t.size;
    if (s == 1 && t.cls != HDM_H5_FLOAT) {  // bytes: copy, test or normalize
      uint8_t *o = (uint8_t *)dst + at;
      if (mode == HDM_H5_AS_U8 && !t.sign) {
        if (o != src) std::memcpy(o, src, n);
        return true;
      }
      if (mode == HDM_H5_AS_NONZERO) {
        for (uint64_t i = 0; i < n; ++i) o[i] = src[i] != 0;
        return true;
      }
      if (mode == HDM_H5_AS_U8) {  // signed bytes: no negatives
        uint8_t neg = 0;
        for (uint64_t i = 0; i < n; ++i) neg |= src[i] & 0x80;
        if (neg) return false;
        if (o != src) std::memcpy(o, src, n);
        return true;
      }
    }
    bool ok = true;
    for (uint64_t i = 0; i < n; ++i) {
      const uint8_t *e = src + i * s;
      double fv = 0;
      int64_t iv = 0;
      bool is_f = t.cls == HDM_H5_FLOAT;
      if (is_f) {
        if (s == 4) {
          uint32_t u;
          std::memcpy(&u, e, 4);
          if (t.be) u = bswap(u);
          float x;
          std::memcpy(&x, &u, 4);
          fv = x;
        } else {
          uint64_t u;
          std::memcpy(&u, e, 8);
          if (t.be) u = bswap(u);
          std::memcpy(&fv, &u, 8);
        }
      } else {
        uint64_t u = 0;
        for (int k = 0; k < s; ++k) u |= (uint64_t)e[t.be ? s - 1 - k : k] << (8 * k);
        if (t.sign && s < 8 && (u >> (8 * s - 1)) & 1) u |= ~0ull << (8 * s);
        if (!t.sign && s == 8 && (u >> 63)) {  // above INT64_MAX
          if (mode == HDM_H5_AS_NONZERO) {
            ((uint8_t *)dst)[at + i] = 1;
            continue;
          }
          ok = false;
          continue;
        }
        iv = (int64_t)u;
      }
      if (mode == HDM_H5_AS_NONZERO) {
        ((uint8_t *)dst)[at + i] = is_f ? (fv != 0.0) : (iv != 0);
        continue;
      }
      if (is_f) {
        if (!(fv == std::floor(fv))) {
          ok = false;
          continue;
        }
        if (fv < -2147483648.0 || fv > 2147483647.0) {
          ok = false;
          continue;
        }
        iv = (int64_t)fv;
      }
      if (mode == HDM_H5_AS_U8) {
        if (iv < 0 || iv > 255) ok = false;
        else ((uint8_t *)dst)[at + i] = (uint8_t)iv;
      } else {
        if (iv < INT32_MIN || iv > INT32_MAX) ok = false;
        else ((int32_t *)dst)[at + i] = (int32_t)iv;
      }
    }
    return ok;
  }
};

uint32_t checksum_fletcher32(const uint8_t *p, size_t n) {
  // Fletcher-32 over 16-bit big-endian words (the HDF5 filter's definition);
  // an odd trailing byte is the high half of a last word
  uint32_t s1 = 0, s2 = 0;
  size_t i = 0;
  while (i + 1 < n) {
    size_t block = std::min<size_t>((n - i) / 2, 360);
    for (size_t k = 0; k < block; ++k, i += 2) {
      s1 += ((uint32_t)p[i] << 8) | p[i + 1];
      s2 += s1;
    }
    s1 = (s1 & 0xffff) + (s1 >> 16);
    s2 = (s2 & 0xffff) + (s2 >> 16);
  }
  if (n & 1) {
    s1 += (uint32_t)p[n - 1] << 8;
    s2 += s1;
    s1 = (s1 & 0xffff) + (s1 >> 16);
    s2 = (s2 & 0xffff) + (s2 >> 16);
  }
  s1 = (s1 & 0xffff) + (s1 >> 16);
  s2 = (s2 & 0xffff) + (s2 >> 16);
  return (s2 << 16) | s1;
}

// Undoes the filter pipeline on one stored chunk (filters applied in order
// when written, undone in reverse; bit i of mask skips filter i).
std::vector<uint8_t> unfilter(const Dataset &d, std::vector<uint8_t> buf, uint32_t mask,
                              uint64_t chunk_bytes) {
  for (int i = (int)d.filters.size() - 1; i >= 0; --i) {
    if ((mask >> i) & 1) continue;
    const Filter &flt = d.filters[i];
    if (flt.id == 3) {  // fletcher32: 4 checksum bytes appended
      if (buf.size() < 4) bad();
      const size_t n = buf.size() - 4;
      uint32_t stored;
      std::memcpy(&stored, buf.data() + n, 4);
      const uint32_t sum = checksum_fletcher32(buf.data(), n);
      // stored little-endian; files of some old library versions hold it
      // byte-reversed
      const uint32_t swapped = __builtin_bswap32(sum);
      if (stored != sum && stored != swapped) bad();
      buf.resize(n);
    } else if (flt.id == 2) {  // shuffle: byte k of every element together
      const size_t es = flt.cd.empty() ? (size_t)d.t.size : flt.cd[0];
      if (es > 1 && buf.size() >= es) {
        std::vector<uint8_t> out(buf.size());
        const size_t ne = buf.size() / es;
        for (size_t k = 0; k < es; ++k)
          for (size_t e = 0; e < ne; ++e) out[e * es + k] = buf[k * ne + e];
        std::memcpy(out.data() + ne * es, buf.data() + ne * es, buf.size() - ne * es);
        buf.swap(out);
      }
    } else if (flt.id == 1) {  // deflate; the output of a later filter's
      // undoing may be longer than the chunk (fletcher32's 4 bytes)
      std::vector<uint8_t> out(chunk_bytes + 64);
      z_stream z{};
      if (inflateInit(&z) != Z_OK) fail(HDM_E_NOMEM);
      z.next_in = buf.data();
      z.avail_in = (uInt)buf.size();
      int r = Z_OK;
      while (r == Z_OK) {
        if (z.total_out == out.size()) {
          if (out.size() > chunk_bytes + 4096) break;  // longer than any filter makes it
          out.resize(out.size() * 2);
        }
        z.next_out = out.data() + z.total_out;
        z.avail_out = (uInt)(out.size() - z.total_out);
        r = inflate(&z, Z_NO_FLUSH);
        if (r == Z_BUF_ERROR && z.avail_out != 0) break;  // input ended early
        if (r == Z_BUF_ERROR) r = Z_OK;
      }
      const uint64_t got = z.total_out;
      inflateEnd(&z);
      if (r != Z_STREAM_END) bad();
      out.resize(got);
      buf.swap(out);
    }
  }
  if (buf.size() < chunk_bytes) bad();
  return buf;
}

struct Chunk {
  uint64_t off[8];  // element offsets of the chunk's first element
  uint64_t addr;
  uint64_t size;  // stored bytes
  uint32_t mask;
};

// chunk records of a B-tree v1 (type 1) index (III.A.1)
void btree_chunks(const File &f, const Dataset &d, uint64_t addr, std::vector<Chunk> &out,
                  uint64_t max_chunks, int depth = 0, int want_level = -1) {
  if (depth > 64) bad();
  f.visit();
  sig(f, addr, "TREE");
  uint8_t h[4];
  f.read(addr + 4, h, 4);
  if (h[0] != 1) bad();
  const int level = h[1];
  if (want_level >= 0 && level != want_level) bad();
  const uint64_t entries = (uint64_t)h[2] | ((uint64_t)h[3] << 8);
  const uint64_t ksz = 8 + 8ull * (d.rank + 1);
  const uint64_t start = addr + 8 + 2ull * f.so;
  const std::vector<uint8_t> body = f.bytes(start, (entries + 1) * ksz + entries * f.so);
  Cur c(body);
  for (uint64_t i = 0; i < entries; ++i) {
    Chunk ch{};
    ch.size = c.u(4);
    ch.mask = (uint32_t)c.u(4);
    for (int k = 0; k <= d.rank; ++k) {
      const uint64_t v = c.u(8);
      if (k < d.rank) ch.off[k] = v;
    }
    ch.addr = c.u(f.so);
    if (level > 0) {
      btree_chunks(f, d, ch.addr, out, max_chunks, depth + 1, level - 1);
    } else {
      if (out.size() >= max_chunks) bad();
      out.push_back(ch);
    }
  }
}

uint64_t chunk_count(const Dataset &d, uint64_t *grid) {
  uint64_t n = 1;
  for (int i = 0; i < d.rank; ++i) {
    grid[i] = d.dims[i] / d.chunk[i] + (d.dims[i] % d.chunk[i] != 0);
    if (__builtin_mul_overflow(n, grid[i], &n)) bad();
  }
  return n;
}

void chunk_origin(const Dataset &d, const uint64_t *grid, uint64_t idx, uint64_t *off) {
  for (int i = d.rank - 1; i >= 0; --i) {
    off[i] = (idx % grid[i]) * d.chunk[i];
    idx /= grid[i];
  }
}

// chunk records of the fixed-array index of layout version 4: header "FAHD"
// then data block "FADB" whose elements are the chunks in row-major order of
// the chunk grid, in pages of 2^page_bits elements when there are more
void fixed_array_chunks(const File &f, const Dataset &d, uint64_t chunk_bytes,
                        std::vector<Chunk> &out) {
  sig(f, d.addr, "FAHD");
  std::vector<uint8_t> hh = f.bytes(d.addr + 4, 4 + f.sl + f.so);
  Cur hc(hh);
  hc.u8();  // version
  const int client = hc.u8();
  const int esize = hc.u8();
  const int page_bits = hc.u8();
  const uint64_t nent = hc.u(f.sl);
  const uint64_t dblk = hc.u(f.so);
  uint64_t grid[8];
  const uint64_t nchunks = chunk_count(d, grid);
  if (nent != nchunks) bad();
  if (dblk == f.undef()) return;  // nothing written: fill value
  sig(f, dblk, "FADB");
  uint64_t pos = dblk + 4 + 2 + f.so;  // signature, version, client id, header address
  const uint64_t page = 1ull << page_bits;
  const bool paged = nent > page;
  const uint64_t npages = paged ? (nent + page - 1) / page : 0;
  std::vector<uint8_t> bitmap;
  if (paged) {  // page-initialized bitmap, then the block's checksum; pages follow
    bitmap = f.bytes(pos, (npages + 7) / 8);
    pos += (npages + 7) / 8 + 4;
  }
  // a filtered entry: address, chunk size (the rest), filter mask (4 bytes)
  const uint64_t fsz = client == 1 ? (uint64_t)esize - f.so - 4 : 0;
  if (client == 1 && (fsz < 1 || fsz > 8)) bad();
  auto parse = [&](const std::vector<uint8_t> &b, uint64_t first, uint64_t count) {
    Cur c(b);
    for (uint64_t i = 0; i < count; ++i) {
      Chunk ch{};
      ch.addr = c.u(f.so);
      if (client == 1) {
        ch.size = c.u((int)fsz);
        ch.mask = (uint32_t)c.u(4);
      } else {
        ch.size = chunk_bytes;
        ch.mask = 0;
      }
      if (ch.addr == f.undef()) continue;
      chunk_origin(d, grid, first + i, ch.off);
      out.push_back(ch);
    }
  };
  if (!paged) {
    parse(f.bytes(pos, nent * esize), 0, nent);
    return;
  }
  for (uint64_t pg = 0; pg < npages; ++pg) {
    const uint64_t first = pg * page;
    const uint64_t cnt = std::min(page, nent - first);
    if ((bitmap[pg / 8] >> (7 - pg % 8)) & 1) parse(f.bytes(pos, cnt * esize), first, cnt);
    pos += cnt * esize + 4;  // each page closes with its checksum
  }
}

struct Job {
  const File *f;
  const Dataset *d;
  Conv cv;
  void *dst;
};

// copies the decoded chunk (chunk dims, C order) into dst, clipped to the
// dataset's extent
bool place_chunk(const Job &j, const Chunk &ch, const uint8_t *data) {
  const Dataset &d = *j.d;
  const int r = d.rank;
  const int es = d.t.size;
  uint64_t ext[8];
  for (int i = 0; i < r; ++i) {
    if (ch.off[i] >= d.dims[i]) return true;  // stray chunk outside the extent
    ext[i] = std::min(d.chunk[i], d.dims[i] - ch.off[i]);
  }
  const uint64_t row = ext[r - 1];
  uint64_t nrows = 1;
  for (int i = 0; i < r - 1; ++i) nrows *= ext[i];
  bool ok = true;
  uint64_t idx[8] = {0};
  for (uint64_t q = 0; q < nrows; ++q) {
    uint64_t rem = q;
    for (int i = r - 2; i >= 0; --i) {
      idx[i] = rem % ext[i];
      rem /= ext[i];
    }
    uint64_t src = 0, dst = 0;
    for (int i = 0; i < r - 1; ++i) {
      src = src * d.chunk[i] + idx[i];
      dst = dst * d.dims[i] + ch.off[i] + idx[i];
    }
    src = src * d.chunk[r - 1];
    dst = dst * d.dims[r - 1] + ch.off[r - 1];
    ok &= j.cv.run(data + src * es, row, j.dst, dst);
  }
  return ok;
}

int hw_threads(int threads) {
  if (threads > 0) return std::min(threads, 64);
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

// runs fn(i) for i in [0, n) on up to `threads` threads; the first error
// code wins
template <typename F>
int parallel_for(uint64_t n, int threads, F fn) {
  std::atomic<uint64_t> next{0};
  std::atomic<int> err{0};
  auto work = [&] {
    for (;;) {
      const uint64_t i = next.fetch_add(1);
      if (i >= n || err.load()) return;
      try {
        const int rc = fn(i);
        if (rc) {
          int z = 0;
          err.compare_exchange_strong(z, rc);
        }
      } catch (const H5Err &e) {
        int z = 0;
        err.compare_exchange_strong(z, e.code);
      } catch (...) {
        int z = 0;
        err.compare_exchange_strong(z, HDM_E_NOMEM);
      }
    }
  };
  const int nt = (int)std::min<uint64_t>((uint64_t)threads, n ? n : 1);
  std::vector<std::thread> ts;
  for (int t = 1; t < nt; ++t) ts.emplace_back(work);
  work();
  for (auto &t : ts) t.join();
  return err.load();
}

int read_dataset(const File &f, const Dataset &d, int mode, void *dst, uint64_t n, int threads) {
  const size_t osz = mode == HDM_H5_AS_I32 ? 4 : 1;
  Job j{&f, &d, Conv{d.t, mode}, dst};
  const uint64_t es = (uint64_t)d.t.size;
  if (d.n == 0) return HDM_OK;
  if (d.layout == 0) {  // compact: the data sits in the layout message
    if (d.compact.size() < d.n * es) bad();
    return j.cv.run(d.compact.data(), d.n, dst, 0) ? HDM_OK : HDM_E_INVALID;
  }
  if (d.layout == 1) {  // contiguous
    if (d.addr == f.undef()) {  // never written: the fill value (0)
      std::memset(dst, 0, n * osz);
      return HDM_OK;
    }
    const uint64_t block = 8ull << 20;  // elements per task
    const uint64_t tasks = (d.n + block - 1) / block;
    return parallel_for(tasks, hw_threads(threads), [&](uint64_t t) -> int {
      const uint64_t a = t * block, cnt = std::min(block, d.n - a);
      if (mode != HDM_H5_AS_I32 && es == 1 && d.t.cls != HDM_H5_FLOAT) {
        uint8_t *o = (uint8_t *)dst + a;  // straight into place, then in place
        f.read(d.addr + a, o, cnt);
        return j.cv.run(o, cnt, dst, a) ? HDM_OK : HDM_E_INVALID;
      }
      std::vector<uint8_t> buf(cnt * es);
      f.read(d.addr + a * es, buf.data(), buf.size());
      return j.cv.run(buf.data(), cnt, dst, a) ? HDM_OK : HDM_E_INVALID;
    });
  }
  // chunked
  uint64_t grid[8];
  const uint64_t nchunks = chunk_count(d, grid);
  uint64_t chunk_elems = 1;
  for (int i = 0; i < d.rank; ++i)
    if (__builtin_mul_overflow(chunk_elems, d.chunk[i], &chunk_elems)) bad();
  if (chunk_elems > (1ull << 32)) bad();  // HDF5 chunks are below 4 GiB
  const uint64_t chunk_bytes = chunk_elems * es;
  std::vector<Chunk> chunks;
  if (d.index == 0) {
    if (d.addr != f.undef()) btree_chunks(f, d, d.addr, chunks, nchunks);
  } else if (d.index == 1) {  // single chunk
    if (d.addr != f.undef()) {
      Chunk ch{};
      ch.addr = d.addr;
      ch.size = d.single_filtered ? d.single_size : chunk_bytes;
      ch.mask = d.single_mask;
      chunks.push_back(ch);
    }
  } else if (d.index == 2) {  // implicit: every chunk, in grid order
    if (chunk_bytes == 0 || nchunks > f.size / chunk_bytes) bad();
    if (d.addr != f.undef())
      for (uint64_t i = 0; i < nchunks; ++i) {
        Chunk ch{};
        ch.addr = d.addr + i * chunk_bytes;
        ch.size = chunk_bytes;
        chunk_origin(d, grid, i, ch.off);
        chunks.push_back(ch);
      }
  } else {
    fixed_array_chunks(f, d, chunk_bytes, chunks);
  }
  if (chunks.size() < nchunks) std::memset(dst, 0, n * osz);  // unwritten chunks: fill 0
  return parallel_for(chunks.size(), hw_threads(threads), [&](uint64_t i) -> int {
    const Chunk &ch = chunks[i];
    std::vector<uint8_t> raw = f.bytes(ch.addr, ch.size);
    const std::vector<uint8_t> data =
        d.filters.empty() ? raw : unfilter(d, std::move(raw), ch.mask, chunk_bytes);
    if (data.size() < chunk_bytes) bad();
    return place_chunk(j, ch, data.data()) ? HDM_OK : HDM_E_INVALID;
  });
}

void fill_info(const Dataset &d, hdm_h5_info *o) {
  std::memset(o, 0, sizeof(*o));
  o->rank = d.rank;
  o->type_class = d.t.cls;
  o->elem_size = d.t.size;
  o->layout = d.layout;
  for (int i = 0; i < d.rank; ++i) o->dims[i] = d.dims[i], o->chunk[i] = d.chunk[i];
  o->n_elems = d.n;
  o->n_filters = (int)d.filters.size();
  o->big_endian = d.t.be;
}

}  // namespace

extern "C" int hdm_h5_info_get(const char *path, const char *dataset, hdm_h5_info *out) {
  if (!out) return HDM_E_INVALID;
  try {
    File f;
    const Dataset d = open_dataset(f, path, dataset);
    fill_info(d, out);
    return HDM_OK;
  } catch (const H5Err &e) {
    return e.code;
  } catch (...) {
    return HDM_E_NOMEM;
  }
}

extern "C" int hdm_h5_read(const char *path, const char *dataset, int conversion, void *dst,
                           uint64_t dst_elems, int threads) {
  if (!dst || (conversion != HDM_H5_AS_U8 && conversion != HDM_H5_AS_I32 &&
               conversion != HDM_H5_AS_NONZERO))
    return HDM_E_INVALID;
  try {
    File f;
    const Dataset d = open_dataset(f, path, dataset);
    if (d.n != dst_elems) return HDM_E_INVALID;
    return read_dataset(f, d, conversion, dst, dst_elems, threads);
  } catch (const H5Err &e) {
    return e.code;
  } catch (...) {
    return HDM_E_NOMEM;
  }
}

extern "C" int hdm_hausdorff_h5(const char *gold_path, const char *gold_dataset,
                                const char *segm_path, const char *segm_dataset, int nx, int ny,
                                int nz, const hdm_opts *opts_in, hdm_result *out,
                                double *load_ms) {
  if (!out || nx < 1 || ny < 1 || nz < 1) return HDM_E_INVALID;
  const uint64_t n = (uint64_t)nx * ny * nz;
  // both volumes are staged in pinned host memory, kept (grow-only) across
  // calls, so they reach the device by DMA; calls serialize on it
  // (pageable memory when pinning fails); both datasets are checked first
  hdm_h5_info info;
  for (int k = 0; k < 2; ++k) {
    const int rc = hdm_h5_info_get(k ? segm_path : gold_path, k ? segm_dataset : gold_dataset,
                                   &info);
    if (rc != HDM_OK) return rc;
    if (info.n_elems != n) return HDM_E_INVALID;
  }
  static std::mutex mu;
  static uint8_t *pinned = nullptr;
  static uint64_t cap = 0;
  std::lock_guard<std::mutex> lk(mu);
  std::vector<uint8_t> pageable;
  uint8_t *g = nullptr;
  if (cap < 2 * n) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr, cap = 0;
    if (hipHostMalloc((void **)&pinned, 2 * n, hipHostMallocDefault) == hipSuccess) {
      cap = 2 * n;
    } else {
      (void)hipGetLastError();
      pinned = nullptr;
    }
  }
  if (pinned) {
    g = pinned;
  } else {
    try {
      pageable.resize(2 * n);
    } catch (...) {
      return HDM_E_NOMEM;
    }
    g = pageable.data();
  }
  uint8_t *s = g + n;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = hdm_h5_read(gold_path, gold_dataset, HDM_H5_AS_NONZERO, g, n, 0);
  if (rc == HDM_OK) rc = hdm_h5_read(segm_path, segm_dataset, HDM_H5_AS_NONZERO, s, n, 0);
  if (load_ms)
    *load_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc != HDM_OK) return rc;
  hdm_opts o;
  hdm_opts_init(&o);
  if (opts_in) o = *opts_in;
  o.ptr_kind = HDM_PTR_HOST;
  o.stream = nullptr;
  return hdm_hausdorff_u8(g, s, nx, ny, nz, 1, &o, out);
}
