// Reading and writing the reference's index file (docs/FORMAT.md, version 1),
// so that a file written by the reference loads here and a file written here
// loads in the reference. The format is a contract, so its constants are the
// reference's: the 256-byte header and field offsets, 64-byte aligned sections
// in the order vectors, level0, labels, levels, deleted, upper, the FNV-1a 64
// header checksum and the section-by-section payload hash
// (src/persist.cpp:84-174), and the load-time validation with its error text
// "corrupt or incompatible index file '<path>': <why>" (src/persist.cpp:241-365).

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "graph.hpp"

namespace hnswamd {

FileMap::FileMap(const std::string& path) {
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) throw std::runtime_error("cannot open " + path);
  struct stat st {};
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    throw std::runtime_error("cannot stat " + path);
  }
  if (st.st_size <= 0) {
    ::close(fd);
    throw std::runtime_error(path + ": empty file");
  }
  const auto n = static_cast<std::size_t>(st.st_size);
  void* p = ::mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (p == MAP_FAILED) throw std::runtime_error("mmap failed for " + path);
  p_ = static_cast<const unsigned char*>(p);
  n_ = n;
}
FileMap::FileMap(FileMap&& o) noexcept : p_(o.p_), n_(o.n_) {
  o.p_ = nullptr;
  o.n_ = 0;
}
FileMap& FileMap::operator=(FileMap&& o) noexcept {
  if (this != &o) {
    if (p_ != nullptr) ::munmap(const_cast<unsigned char*>(p_), n_);
    p_ = o.p_;
    n_ = o.n_;
    o.p_ = nullptr;
    o.n_ = 0;
  }
  return *this;
}
FileMap::~FileMap() {
  if (p_ != nullptr) ::munmap(const_cast<unsigned char*>(p_), n_);
}

namespace {

constexpr std::size_t kHeaderBytes = 256;
constexpr uint32_t kFormatVersion = 1;
constexpr uint32_t kEndianTag = 0x01020304u;
const char kMagicBytes[8] = {'H', 'N', 'S', 'W', 'E', 'N', 'G', '\0'};

// Header field offsets (docs/FORMAT.md, "Header").
enum HeaderOffset : std::size_t {
  kOffMagic = 0, kOffVersion = 8, kOffEndian = 12, kOffDim = 16, kOffMetric = 24, kOffM = 32, kOffEfc = 40,
  kOffEfSearch = 48, kOffSeed = 56, kOffMaxElements = 64, kOffCount = 72, kOffPadded = 80, kOffMaxLevel = 88,
  kOffEntry = 92, kOffUpperUsed = 96, kOffNumDeleted = 104, kOffRng = 112, kOffSecVectors = 120,
  kOffSecLevel0 = 128, kOffSecLabels = 136, kOffSecLevels = 144, kOffSecDeleted = 152, kOffSecUpper = 160,
  kOffFileSize = 168, kOffPayloadSum = 176, kOffHeaderSum = 184
};

template <typename T>
void store(unsigned char* b, std::size_t at, T v) {
  std::memcpy(b + at, &v, sizeof(T));
}
template <typename T>
T fetch(const unsigned char* b, std::size_t at) {
  T v;
  std::memcpy(&v, b + at, sizeof(T));
  return v;
}

uint64_t header_fnv(const unsigned char* hdr) {
  uint64_t h = 0xcbf29ce484222325ULL;
  for (std::size_t i = 0; i < kHeaderBytes; ++i) {
    const unsigned char c = (i >= kOffHeaderSum && i < kOffHeaderSum + 8) ? 0 : hdr[i];
    h = (h ^ c) * 0x100000001b3ULL;
  }
  return h;
}

class SectionHash {
 public:
  void feed(const void* data, std::size_t n) {
    const auto* p = static_cast<const unsigned char*>(data);
    std::size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, p + i, 8);
      mix(w);
    }
    for (; i < n; ++i) mix(0x100ULL | p[i]);
    mix(n);
  }
  uint64_t value() const { return h_; }

 private:
  void mix(uint64_t w) {
    h_ ^= w * 0x9E3779B97F4A7C15ULL;
    h_ = (h_ << 31) | (h_ >> 33);
    h_ *= 0xC2B2AE3D27D4EB4FULL;
  }
  uint64_t h_ = 0x243F6A8885A308D3ULL;
};

struct Fields {
  uint64_t dim = 0, M = 0, efc = 0, ef_search = 0, seed = 0, max_elements = 0, count = 0, padded = 0;
  uint32_t metric = 0, entry = kNoNode;
  int32_t max_level = -1;
  uint64_t upper_used = 0, num_deleted = 0, rng = 0;
  uint64_t sec[6] = {};  // vectors, level0, labels, levels, deleted, upper
  uint64_t file_size = 0, payload_sum = 0;
};

void section_sizes(const Fields& f, uint64_t out[6]) {
  out[0] = f.count * f.padded * sizeof(float);
  out[1] = f.count * (1 + 2 * f.M) * sizeof(uint32_t);
  out[2] = f.count * sizeof(uint64_t);
  out[3] = f.count;
  out[4] = f.count;
  out[5] = f.upper_used * sizeof(uint32_t);
}

void place_sections(Fields& f) {
  uint64_t sz[6];
  section_sizes(f, sz);
  uint64_t pos = kHeaderBytes;
  for (int s = 0; s < 6; ++s) {
    f.sec[s] = pos;
    pos += sz[s];
    if (s < 5) pos = round_up(pos, kRowAlign);
  }
  f.file_size = pos;
}

[[noreturn]] void bad_file(const std::string& path, const std::string& why) {
  throw std::runtime_error("corrupt or incompatible index file '" + path + "': " + why);
}

void encode(const Fields& f, unsigned char* b) {
  std::memset(b, 0, kHeaderBytes);
  std::memcpy(b + kOffMagic, kMagicBytes, 8);
  store<uint32_t>(b, kOffVersion, kFormatVersion);
  store<uint32_t>(b, kOffEndian, kEndianTag);
  store(b, kOffDim, f.dim);
  store(b, kOffMetric, f.metric);
  store(b, kOffM, f.M);
  store(b, kOffEfc, f.efc);
  store(b, kOffEfSearch, f.ef_search);
  store(b, kOffSeed, f.seed);
  store(b, kOffMaxElements, f.max_elements);
  store(b, kOffCount, f.count);
  store(b, kOffPadded, f.padded);
  store(b, kOffMaxLevel, f.max_level);
  store(b, kOffEntry, f.entry);
  store(b, kOffUpperUsed, f.upper_used);
  store(b, kOffNumDeleted, f.num_deleted);
  store(b, kOffRng, f.rng);
  for (int s = 0; s < 6; ++s) store(b, kOffSecVectors + 8 * static_cast<std::size_t>(s), f.sec[s]);
  store(b, kOffFileSize, f.file_size);
  store(b, kOffPayloadSum, f.payload_sum);
  store<uint64_t>(b, kOffHeaderSum, header_fnv(b));
}

Fields decode(const unsigned char* b, std::size_t actual, const std::string& path) {
  if (actual < kHeaderBytes) bad_file(path, "file too small for header");
  if (std::memcmp(b + kOffMagic, kMagicBytes, 8) != 0) bad_file(path, "bad magic");
  if (fetch<uint32_t>(b, kOffEndian) != kEndianTag) bad_file(path, "endianness mismatch");
  if (fetch<uint32_t>(b, kOffVersion) != kFormatVersion) {
    bad_file(path, "unsupported version " + std::to_string(fetch<uint32_t>(b, kOffVersion)));
  }
  if (fetch<uint64_t>(b, kOffHeaderSum) != header_fnv(b)) bad_file(path, "header checksum mismatch");
  Fields f;
  f.dim = fetch<uint64_t>(b, kOffDim);
  f.metric = fetch<uint32_t>(b, kOffMetric);
  f.M = fetch<uint64_t>(b, kOffM);
  f.efc = fetch<uint64_t>(b, kOffEfc);
  f.ef_search = fetch<uint64_t>(b, kOffEfSearch);
  f.seed = fetch<uint64_t>(b, kOffSeed);
  f.max_elements = fetch<uint64_t>(b, kOffMaxElements);
  f.count = fetch<uint64_t>(b, kOffCount);
  f.padded = fetch<uint64_t>(b, kOffPadded);
  f.max_level = fetch<int32_t>(b, kOffMaxLevel);
  f.entry = fetch<uint32_t>(b, kOffEntry);
  f.upper_used = fetch<uint64_t>(b, kOffUpperUsed);
  f.num_deleted = fetch<uint64_t>(b, kOffNumDeleted);
  f.rng = fetch<uint64_t>(b, kOffRng);
  for (int s = 0; s < 6; ++s) f.sec[s] = fetch<uint64_t>(b, kOffSecVectors + 8 * static_cast<std::size_t>(s));
  f.file_size = fetch<uint64_t>(b, kOffFileSize);
  f.payload_sum = fetch<uint64_t>(b, kOffPayloadSum);

  if (f.file_size != actual) bad_file(path, "size mismatch (truncated or extended file)");
  if (f.dim == 0 || f.dim > (1u << 20)) bad_file(path, "invalid dim");
  if (f.metric > 2) bad_file(path, "invalid metric");
  if (f.M < 2 || f.M > 2048) bad_file(path, "invalid M");
  if (f.efc == 0 || f.ef_search == 0) bad_file(path, "invalid ef");
  if (f.padded != round_up(f.dim, kPadFloats)) bad_file(path, "invalid padded dim");
  if (f.count > actual || f.count >= kNoNode) bad_file(path, "invalid count");
  if (f.upper_used > actual) bad_file(path, "invalid upper size");
  if (f.num_deleted > f.count) bad_file(path, "invalid deleted count");
  if (f.max_elements != 0 && (f.max_elements < f.count || f.max_elements >= kNoNode)) {
    bad_file(path, "invalid max_elements");
  }
  if (f.count == 0) {
    if (f.max_level != -1 || f.entry != kNoNode) bad_file(path, "invalid empty-index header");
  } else if (f.max_level < 0 || f.max_level > kMaxLevel || f.entry >= f.count) {
    bad_file(path, "invalid entry point / max level");
  }
  Fields expect = f;
  place_sections(expect);
  for (int s = 0; s < 6; ++s) {
    if (expect.sec[s] != f.sec[s]) bad_file(path, "section offsets inconsistent with header");
  }
  if (expect.file_size != f.file_size) bad_file(path, "section offsets inconsistent with header");
  return f;
}

BuildParams params_of(const Fields& f) {
  BuildParams p;
  p.dim = f.dim;
  p.metric = static_cast<int>(f.metric);
  p.M = f.M;
  p.ef_construction = f.efc;
  p.max_elements = 0;  // storage is sized from count below; the cap is restored after
  p.seed = f.seed;
  return p;
}

/// Graph checks and derived arrays (upper offsets, label map, tombstones).
void index_graph(Graph& g, const Fields& f, const uint8_t* del, const std::string& path) {
  const std::size_t n = f.count;
  g.upper_offset.assign(std::max(n, g.capacity), 0);
  uint64_t pos = 0;
  for (std::size_t i = 0; i < n; ++i) {
    const int lv = g.levels[i];
    if (lv > f.max_level) bad_file(path, "node level exceeds max level");
    g.upper_offset[i] = pos;
    pos += static_cast<uint64_t>(lv) * g.upper_stride;
  }
  if (pos != f.upper_used) bad_file(path, "upper-layer size does not match levels");
  if (n > 0 && g.levels[f.entry] != f.max_level) bad_file(path, "entry point is not on the top level");
  for (std::size_t i = 0; i < n; ++i) {
    const auto id = static_cast<uint32_t>(i);
    for (int lc = 0; lc <= g.levels[i]; ++lc) {
      const uint32_t* l = g.list(id, lc);
      if (l[0] > g.degree_cap(lc)) bad_file(path, "neighbor count exceeds maximum degree");
      for (uint32_t j = 0; j < l[0]; ++j) {
        const uint32_t nb = l[1 + j];
        if (nb >= n || g.levels[nb] < lc) bad_file(path, "neighbor id out of range");
      }
    }
  }
  g.label_to_id.clear();
  g.label_to_id.reserve(n);
  for (std::size_t i = 0; i < n; ++i) {
    if (!g.label_to_id.emplace(g.labels[i], static_cast<uint32_t>(i)).second) bad_file(path, "duplicate label");
  }
  g.deleted.assign(del, del + n);
  std::size_t dels = 0;
  for (uint8_t d : g.deleted) {
    if (d > 1) bad_file(path, "invalid tombstone byte");
    dels += d;
  }
  if (dels != f.num_deleted) bad_file(path, "deleted count mismatch");
  g.num_deleted = dels;
}

}  // namespace

void Graph::save(const std::string& path) const {
  Fields f;
  f.dim = dim;
  f.metric = static_cast<uint32_t>(params.metric);
  f.M = M;
  f.efc = params.ef_construction;
  f.ef_search = ef_search;
  f.seed = params.seed;
  f.max_elements = params.max_elements;
  f.count = count;
  f.padded = padded_dim;
  f.max_level = max_level;
  f.entry = entry_point;
  f.upper_used = upper_used;
  f.num_deleted = num_deleted;
  f.rng = rng_state;
  place_sections(f);
  uint64_t sz[6];
  section_sizes(f, sz);
  const void* src[6] = {vecs, l0, labels, levels, deleted.data(), upper};
  SectionHash hash;
  for (int s = 0; s < 6; ++s) hash.feed(src[s], sz[s]);
  f.payload_sum = hash.value();

  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("cannot open " + path + " for writing");
  unsigned char hdr[kHeaderBytes];
  encode(f, hdr);
  out.write(reinterpret_cast<const char*>(hdr), kHeaderBytes);
  const char zeros[kRowAlign] = {};
  uint64_t pos = kHeaderBytes;
  for (int s = 0; s < 6; ++s) {
    out.write(zeros, static_cast<std::streamsize>(f.sec[s] - pos));
    if (sz[s] > 0) out.write(static_cast<const char*>(src[s]), static_cast<std::streamsize>(sz[s]));
    pos = f.sec[s] + sz[s];
  }
  out.flush();
  if (!out) throw std::runtime_error("write failed: " + path);
}

std::unique_ptr<Graph> Graph::load(const std::string& path, bool mmap) {
  if (mmap) {
    FileMap file(path);
    const Fields f = decode(file.data(), file.size(), path);
    auto g = std::make_unique<Graph>(params_of(f));
    g->params.max_elements = f.max_elements;
    g->count = g->capacity = f.count;
    g->ef_search = f.ef_search;
    g->upper_used = f.upper_used;
    g->entry_point = f.entry;
    g->max_level = f.max_level;
    g->rng_state = f.rng;
    const unsigned char* b = file.data();
    g->vecs = reinterpret_cast<const float*>(b + f.sec[0]);
    g->l0 = const_cast<uint32_t*>(reinterpret_cast<const uint32_t*>(b + f.sec[1]));
    g->labels = reinterpret_cast<const uint64_t*>(b + f.sec[2]);
    g->levels = b + f.sec[3];
    g->upper = const_cast<uint32_t*>(reinterpret_cast<const uint32_t*>(b + f.sec[5]));
    uint64_t sz[6];
    section_sizes(f, sz);
    SectionHash hash;
    uint64_t pos = kHeaderBytes;
    for (int s = 0; s < 6; ++s) {
      for (uint64_t i = pos; i < f.sec[s]; ++i) {
        if (b[i] != 0) bad_file(path, "non-zero padding");
      }
      hash.feed(b + f.sec[s], sz[s]);
      pos = f.sec[s] + sz[s];
    }
    if (hash.value() != f.payload_sum) bad_file(path, "payload checksum mismatch");
    index_graph(*g, f, b + f.sec[4], path);
    g->mapping = std::move(file);
    g->read_only = true;
    return g;
  }

  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error("cannot open " + path);
  const auto actual = static_cast<std::size_t>(in.tellg());
  unsigned char hdr[kHeaderBytes];
  if (actual < kHeaderBytes) bad_file(path, "file too small for header");
  in.seekg(0);
  if (!in.read(reinterpret_cast<char*>(hdr), kHeaderBytes)) bad_file(path, "short read");
  const Fields f = decode(hdr, actual, path);
  auto g = std::make_unique<Graph>(params_of(f));
  g->params.max_elements = f.max_elements;
  if (f.count > 0) g->reserve(std::max<std::size_t>(f.count, f.max_elements));
  uint64_t sz[6];
  section_sizes(f, sz);
  std::vector<uint8_t> del(f.count);
  g->upper_store.assign(f.upper_used, 0u);
  void* dst[6] = {g->vec_store.data(), g->l0_store.data(), g->label_store.data(), g->level_store.data(),
                  del.data(), g->upper_store.data()};
  SectionHash hash;
  uint64_t pos = kHeaderBytes;
  for (int s = 0; s < 6; ++s) {
    const uint64_t gap = f.sec[s] - pos;
    char pad[kRowAlign];
    if (gap > sizeof(pad) || (gap > 0 && !in.read(pad, static_cast<std::streamsize>(gap)))) {
      bad_file(path, "short read");
    }
    for (uint64_t i = 0; i < gap; ++i) {
      if (pad[i] != 0) bad_file(path, "non-zero padding");
    }
    if (sz[s] > 0 && !in.read(static_cast<char*>(dst[s]), static_cast<std::streamsize>(sz[s]))) {
      bad_file(path, "short read");
    }
    hash.feed(dst[s], sz[s]);
    pos = f.sec[s] + sz[s];
  }
  if (hash.value() != f.payload_sum) bad_file(path, "payload checksum mismatch");
  g->rebind();
  g->count = f.count;
  g->ef_search = f.ef_search;
  g->upper_used = f.upper_used;
  g->entry_point = f.entry;
  g->max_level = f.max_level;
  g->rng_state = f.rng;
  index_graph(*g, f, del.data(), path);
  g->deleted.resize(g->capacity, 0);
  return g;
}

}  // namespace hnswamd
