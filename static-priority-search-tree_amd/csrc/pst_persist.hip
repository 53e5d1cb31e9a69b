// Tree persistence and the raw volume reader.
//
// A saved tree is a 64-byte header and the node records as SoA sections,
// one per field, each over all S record slots in slot order:
//
//   header  magic "PSTAMD", version, kind (0 StaticPSTGPU, 1 StaticPSTGPUArr),
//           dtype, id_dtype, n, S, slots_per_tree, checksum
//   dim1[S] dim2[S] med[S] (sizeof(T) each), then id[S] (sizeof(IDType) each,
//   absent without IDs)
//
// Slots that hold no point (the unused end of the last level, the forest's
// padding slot per tree) are written as zeros, so a file depends only on
// the tree. The checksum is a four-lane 64-bit multiply-xor hash over each
// section's bytes, combined in section order. Loading validates the header
// against the file size, rebuilds the records from the sections, and
// recomputes the per-slot side arrays (subtree minima, IDs).
//
// The reference has no on-disk form for its trees; it only prints them
// (pst/static-pst-gpu.tu:441-483).
//
// readInVertices (helpers/isosurface-data-processing.h:205-229) reads
// num_vertices values of T from a raw binary file; read_raw streams the file
// to host or, through pinned staging, to device memory, with the
// reference's two error messages.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr size_t kChunkSlots = size_t{1} << 22;
constexpr int kErrIo = 6;  // the C ABI's PST_ERR_IO
constexpr char kMagic[8] = {'P', 'S', 'T', 'A', 'M', 'D', '\0', '\1'};
constexpr uint32_t kVersion = 1;

// Hash of a byte stream fed in pieces whose sizes are multiples of 8 (the
// last piece may end with a partial word, padded with zeros).
struct Hasher {
  uint64_t h[4] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull, 0x27D4EB2F165667C5ull};
  uint64_t words = 0;
  void feed(const unsigned char* p, size_t bytes) {
    const size_t full = bytes / 8;
    for (size_t i = 0; i < full; i++) {
      uint64_t w;
      std::memcpy(&w, p + 8 * i, 8);
      uint64_t& x = h[(words + i) & 3];
      x = (x ^ w) * 0x100000001B3ull;
      x ^= x >> 29;
    }
    words += full;
    if (bytes % 8) {
      uint64_t w = 0;
      std::memcpy(&w, p + 8 * full, bytes % 8);
      uint64_t& x = h[words & 3];
      x = (x ^ w) * 0x100000001B3ull;
      x ^= x >> 29;
      words++;
    }
  }
  uint64_t digest() const {
    uint64_t d = words;
    for (int i = 0; i < 4; i++) d = (d ^ h[i]) * 0xff51afd7ed558ccdull, d ^= d >> 33;
    return d;
  }
};

// Slot geometry: slot s is in tree s >> shift at local slot s - (tree <<
// shift); local slots at or beyond `per_tree` are padding. One deep tree:
// shift 31, one tree, per_tree = S.
struct Geometry {
  uint32_t shift, ntrees, full_M, last_M;
  uint64_t per_tree;
};

Geometry geometry(size_t n, size_t slots_per_tree, size_t S) {
  Geometry g{};
  if (slots_per_tree == 0) {
    g.shift = 31;
    g.ntrees = 1;
    g.full_M = g.last_M = static_cast<uint32_t>(n + 1);
    g.per_tree = S;
  } else {
    g.shift = 0;
    while ((size_t{1} << g.shift) < slots_per_tree + 1) g.shift++;
    g.ntrees = static_cast<uint32_t>((n + slots_per_tree - 1) / slots_per_tree);
    g.full_M = static_cast<uint32_t>(slots_per_tree + 1);
    g.last_M = static_cast<uint32_t>(n - (g.ntrees - 1ull) * slots_per_tree + 1);
    g.per_tree = slots_per_tree;
  }
  return g;
}

__device__ inline bool slot_holds_point(uint64_t s, const Geometry& g) {
  const uint32_t tree = static_cast<uint32_t>(s >> g.shift);
  const uint64_t local = s - (static_cast<uint64_t>(tree) << g.shift);
  if (local >= g.per_tree || tree >= g.ntrees) return false;
  return subtree_size32(tree + 1 < g.ntrees ? g.full_M : g.last_M, static_cast<uint32_t>(local)) >= 1;
}

// records [base, base + cnt) -> sections of cnt entries each
template <typename T, typename IS>
__global__ void k_to_soa(const Node<T, IS>* __restrict__ nodes, uint64_t base, uint64_t cnt, Geometry g,
                         T* __restrict__ d1, T* __restrict__ d2, T* __restrict__ md, IS* __restrict__ id) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const bool live = slot_holds_point(base + i, g);
  Node<T, IS> nd{};
  if (live) nd = nodes[base + i];
  d1[i] = live ? nd.dim1 : T{};
  d2[i] = live ? nd.dim2 : T{};
  md[i] = live ? nd.med : T{};
  if constexpr (has_id_v<IS>) id[i] = live ? nd.id : IS{};
}

template <typename T, typename IS>
__global__ void k_from_soa(Node<T, IS>* __restrict__ nodes, uint64_t base, uint64_t cnt, const T* __restrict__ d1,
                           const T* __restrict__ d2, const T* __restrict__ md, const IS* __restrict__ id) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  Node<T, IS> nd{};
  nd.dim1 = d1[i];
  nd.dim2 = d2[i];
  nd.med = md[i];
  if constexpr (has_id_v<IS>) nd.id = id[i];
  nodes[base + i] = nd;
}

template <typename F>
Status with_t(int dt, F&& f) {
  switch (dt) {
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
  }
  return Status{1, "Error: unsupported data type code " + std::to_string(dt)};
}
template <typename F>
Status with_is(int idt, F&& f) {
  switch (idt) {
    case NO_ID: return f(NoId{});
    case F32: case I32: case U32: return f(uint32_t{});
    case F64: case I64: case U64: return f(uint64_t{});
  }
  return Status{1, "Error: unsupported ID type code " + std::to_string(idt)};
}

size_t t_size(int dt) { return (dt == F64 || dt == I64) ? 8 : 4; }
size_t i_size(int idt) {
  if (idt == NO_ID) return 0;
  return (idt == F64 || idt == I64 || idt == U64) ? 8 : 4;
}

struct Header {
  char magic[8];
  uint32_t version;
  uint32_t kind;
  int32_t dtype, id_dtype;
  uint64_t n, S, slots_per_tree, checksum;
};
static_assert(sizeof(Header) == 56, "header layout");
constexpr size_t kHeaderBytes = 64;

// Pinned host staging and device staging for one chunk of slots.
struct Staging {
  void* host = nullptr;
  void* dev = nullptr;
  ~Staging() {
    if (host) hipHostFree(host);
    if (dev) hipFree(dev);
  }
  Status alloc(size_t bytes) {
    hipError_t e = hipHostMalloc(&host, bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&dev, bytes);
    return hip_status(e, "Error in allocating tree file staging buffers");
  }
};

bool write_all(int fd, const void* p, size_t bytes, off_t off) {
  const char* c = static_cast<const char*>(p);
  while (bytes) {
    const ssize_t w = pwrite(fd, c, bytes, off);
    if (w <= 0) return false;
    c += w;
    bytes -= static_cast<size_t>(w);
    off += w;
  }
  return true;
}
bool read_all(int fd, void* p, size_t bytes, off_t off) {
  char* c = static_cast<char*>(p);
  while (bytes) {
    const ssize_t r = pread(fd, c, bytes, off);
    if (r <= 0) return false;
    c += r;
    bytes -= static_cast<size_t>(r);
    off += r;
  }
  return true;
}

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

}  // namespace

Status save_tree_file(const char* path, int kind, int dtype, int id_dtype, size_t n, size_t S,
                      size_t slots_per_tree, const void* nodes_d, hipStream_t stream) {
  const size_t sT = t_size(dtype), sI = i_size(id_dtype);
  Fd f;
  f.fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (f.fd < 0) return Status{kErrIo, std::string("Failed to open ") + path + " for writing"};
  Header h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.kind = static_cast<uint32_t>(kind);
  h.dtype = dtype;
  h.id_dtype = id_dtype;
  h.n = n;
  h.S = S;
  h.slots_per_tree = slots_per_tree;
  Hasher hs[4];
  if (S > 0) {
    const Geometry g = geometry(n, slots_per_tree, S);
    const size_t per_slot = 3 * sT + sI;
    Staging st;
    Status s0 = st.alloc(kChunkSlots * per_slot);
    if (!s0.ok()) return s0;
    for (size_t base = 0; base < S; base += kChunkSlots) {
      const size_t cnt = std::min(kChunkSlots, S - base);
      char* dv = static_cast<char*>(st.dev);
      Status sk = with_t(dtype, [&](auto t) {
        return with_is(id_dtype, [&](auto i) {
          using T = decltype(t);
          using IS = decltype(i);
          T* d1 = reinterpret_cast<T*>(dv);
          IS* id = reinterpret_cast<IS*>(dv + 3 * cnt * sT);
          k_to_soa<T, IS><<<static_cast<unsigned>((cnt + 255) / 256), 256, 0, stream>>>(
              static_cast<const Node<T, IS>*>(nodes_d), base, cnt, g, d1, d1 + cnt, d1 + 2 * cnt, id);
          return hip_status(hipGetLastError(), "Error in launching tree export for saving");
        });
      });
      if (!sk.ok()) return sk;
      hipError_t e = hipMemcpyAsync(st.host, st.dev, cnt * per_slot, hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return hip_status(e, "Error in copying tree records for saving");
      const unsigned char* hp = static_cast<const unsigned char*>(st.host);
      const size_t fsz[4] = {sT, sT, sT, sI};
      size_t off_in = 0, sec_off = kHeaderBytes;
      for (int k = 0; k < 4; k++) {
        if (fsz[k] == 0) continue;
        const size_t bytes = cnt * fsz[k];
        hs[k].feed(hp + off_in, bytes);
        if (!write_all(f.fd, hp + off_in, bytes, static_cast<off_t>(sec_off + base * fsz[k])))
          return Status{kErrIo, std::string("Failure: could not write ") + path};
        off_in += bytes;
        sec_off += S * fsz[k];
      }
    }
  }
  uint64_t c = 0x6A09E667F3BCC909ull;
  for (int k = 0; k < 4; k++) c = (c ^ hs[k].digest()) * 0x100000001B3ull, c ^= c >> 31;
  h.checksum = c;
  unsigned char hb[kHeaderBytes] = {};
  std::memcpy(hb, &h, sizeof(h));
  if (!write_all(f.fd, hb, kHeaderBytes, 0)) return Status{kErrIo, std::string("Failure: could not write ") + path};
  return Status{};
}

Status read_tree_header(const char* path, TreeFileInfo* info) {
  Fd f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) return Status{kErrIo, std::string("Failed to open ") + path + " for reading"};
  unsigned char hb[kHeaderBytes];
  if (!read_all(f.fd, hb, kHeaderBytes, 0)) return Status{kErrIo, std::string("Error: ") + path + " is not a tree file"};
  Header h;
  std::memcpy(&h, hb, sizeof(h));
  if (std::memcmp(h.magic, kMagic, 8) != 0) return Status{kErrIo, std::string("Error: ") + path + " is not a tree file"};
  if (h.version != kVersion)
    return Status{kErrIo, "Error: tree file version " + std::to_string(h.version) + " is not supported"};
  const bool t_ok = h.dtype >= F32 && h.dtype <= I64;
  const bool i_ok = h.id_dtype == NO_ID || (h.id_dtype >= F32 && h.id_dtype <= U64);
  if (!t_ok || !i_ok || h.kind > 1 || (h.kind == 0) != (h.slots_per_tree == 0))
    return Status{kErrIo, std::string("Error: corrupt header in ") + path};
  // the slot count follows from n and the tree kind
  size_t S = 0;
  if (h.n > 0) {
    if (h.kind == 0) {
      if (h.n >= (1ull << 30)) return Status{kErrIo, std::string("Error: corrupt header in ") + path};
      size_t p = 1;
      while (p <= h.n) p <<= 1;
      S = p - 1;
    } else {
      const uint64_t spt = h.slots_per_tree;
      if (spt == 0 || spt > 2047 || ((spt + 1) & spt) != 0 || h.n >= (1ull << 29))
        return Status{kErrIo, std::string("Error: corrupt header in ") + path};
      S = ((h.n + spt - 1) / spt) * (spt + 1);
    }
  }
  if (S != h.S) return Status{kErrIo, std::string("Error: corrupt header in ") + path};
  struct stat sb;
  const size_t want = kHeaderBytes + S * (3 * t_size(h.dtype) + i_size(h.id_dtype));
  const size_t have = fstat(f.fd, &sb) == 0 ? static_cast<size_t>(sb.st_size) : 0;
  if (have != want)
    return Status{kErrIo, std::string("Error: ") + path + " holds " + std::to_string(have) + " bytes, expected " +
                              std::to_string(want)};
  info->kind = static_cast<int>(h.kind);
  info->dtype = h.dtype;
  info->id_dtype = h.id_dtype;
  info->n = h.n;
  info->S = h.S;
  info->slots_per_tree = h.slots_per_tree;
  info->checksum = h.checksum;
  return Status{};
}

Status load_tree_file(const char* path, const TreeFileInfo& info, void* nodes_d, hipStream_t stream) {
  const size_t S = info.S;
  if (S == 0) return Status{};
  const size_t sT = t_size(info.dtype), sI = i_size(info.id_dtype);
  Fd f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) return Status{kErrIo, std::string("Failed to open ") + path + " for reading"};
  const size_t per_slot = 3 * sT + sI;
  Staging st;
  Status s0 = st.alloc(kChunkSlots * per_slot);
  if (!s0.ok()) return s0;
  Hasher hs[4];
  const size_t fsz[4] = {sT, sT, sT, sI};
  for (size_t base = 0; base < S; base += kChunkSlots) {
    const size_t cnt = std::min(kChunkSlots, S - base);
    unsigned char* hp = static_cast<unsigned char*>(st.host);
    size_t off_in = 0, sec_off = kHeaderBytes;
    for (int k = 0; k < 4; k++) {
      if (fsz[k] == 0) continue;
      const size_t bytes = cnt * fsz[k];
      if (!read_all(f.fd, hp + off_in, bytes, static_cast<off_t>(sec_off + base * fsz[k])))
        return Status{kErrIo, std::string("Failure: could not read ") + path};
      hs[k].feed(hp + off_in, bytes);
      off_in += bytes;
      sec_off += S * fsz[k];
    }
    hipError_t e = hipMemcpyAsync(st.dev, st.host, cnt * per_slot, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_status(e, "Error in copying tree records from a file");
    char* dv = static_cast<char*>(st.dev);
    Status sk = with_t(info.dtype, [&](auto t) {
      return with_is(info.id_dtype, [&](auto i) {
        using T = decltype(t);
        using IS = decltype(i);
        const T* d1 = reinterpret_cast<const T*>(dv);
        const IS* id = reinterpret_cast<const IS*>(dv + 3 * cnt * sT);
        k_from_soa<T, IS><<<static_cast<unsigned>((cnt + 255) / 256), 256, 0, stream>>>(
            static_cast<Node<T, IS>*>(nodes_d), base, cnt, d1, d1 + cnt, d1 + 2 * cnt, id);
        return hip_status(hipGetLastError(), "Error in launching tree import");
      });
    });
    if (!sk.ok()) return sk;
    e = hipStreamSynchronize(stream);  // the staging buffers are reused by the next chunk
    if (e != hipSuccess) return hip_status(e, "Error in loading tree records");
  }
  uint64_t c = 0x6A09E667F3BCC909ull;
  for (int k = 0; k < 4; k++) c = (c ^ hs[k].digest()) * 0x100000001B3ull, c ^= c >> 31;
  if (c != info.checksum) return Status{kErrIo, std::string("Error: checksum mismatch in ") + path};
  return Status{};
}

Status read_raw(const char* path, size_t bytes, void* host_out, void* dev_out, hipStream_t stream) {
  Fd f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) return Status{kErrIo, std::string("Failed to open ") + path + " for reading"};
  struct stat sb;
  const size_t have = fstat(f.fd, &sb) == 0 ? static_cast<size_t>(sb.st_size) : 0;
  if (have < bytes)
    return Status{kErrIo, "Failure: read " + std::to_string(have) + " bytes, expected " + std::to_string(bytes) + " bytes"};
  if (host_out) {
    if (!read_all(f.fd, host_out, bytes, 0))
      return Status{kErrIo, "Failure: read fewer than " + std::to_string(bytes) + " bytes of " + path};
    return Status{};
  }
  // to device: two pinned chunks, so that reading one overlaps copying the other
  constexpr size_t kRawChunk = size_t{64} << 20;
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  hipError_t e = hipSuccess;
  for (int k = 0; k < 2 && e == hipSuccess; k++) {
    e = hipHostMalloc(&pin[k], kRawChunk, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&done[k], hipEventDisableTiming);
  }
  Status st = hip_status(e, "Error in allocating read staging buffers");
  bool used[2] = {false, false};
  for (size_t off = 0, k = 0; st.ok() && off < bytes; off += kRawChunk, k ^= 1) {
    const size_t cnt = std::min(kRawChunk, bytes - off);
    if (used[k]) {
      e = hipEventSynchronize(done[k]);
      if (e != hipSuccess) {
        st = hip_status(e, "Error in copying a volume to the device");
        break;
      }
    }
    if (!read_all(f.fd, pin[k], cnt, static_cast<off_t>(off))) {
      st = Status{kErrIo, "Failure: read fewer than " + std::to_string(bytes) + " bytes of " + path};
      break;
    }
    e = hipMemcpyAsync(static_cast<char*>(dev_out) + off, pin[k], cnt, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(done[k], stream);
    if (e != hipSuccess) st = hip_status(e, "Error in copying a volume to the device");
    used[k] = true;
  }
  const hipError_t es = hipStreamSynchronize(stream);
  if (st.ok() && es != hipSuccess) st = hip_status(es, "Error in copying a volume to the device");
  for (int k = 0; k < 2; k++) {
    if (done[k]) hipEventDestroy(done[k]);
    if (pin[k]) hipHostFree(pin[k]);
  }
  return st;
}

}  // namespace pst
