// On-disk problem format (include/emagddm_problem.h): writer, validating reader.
//
// The reference declares fe_em::MeshReader::read (include/fe_em/mesh_reader.hpp:32-56)
// with no implementation; this is the storage side of one.  Host code only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "emagddm_problem.h"

namespace {

thread_local std::string g_err;

constexpr char kMagic[8] = {'E', 'M', 'A', 'G', 'F', 'D', 'P', '\0'};

// FNV-1a over 64-bit little-endian words (every written field is 8-byte aligned).
struct Hash {
  uint64_t h = 0xcbf29ce484222325ull;
  void words(const void* p, size_t bytes) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < bytes; i += 8) {
      uint64_t w = 0;
      std::memcpy(&w, c + i, bytes - i >= 8 ? 8 : bytes - i);
      h = (h ^ w) * 0x100000001b3ull;
    }
  }
};

struct Header {
  uint32_t version, flags;
  int64_t n_nodes, n_tets, n_pec;
  int32_t n_mat, n_sub;
};
static_assert(sizeof(Header) == 40, "header layout");

struct Writer {
  FILE* f;
  Hash hash;
  bool ok = true;
  void put(const void* p, size_t bytes) {
    if (bytes == 0) return;
    ok = ok && std::fwrite(p, 1, bytes, f) == bytes;
    hash.words(p, bytes);
  }
  // an array padded with zeros to a multiple of 8 bytes; the hash sees the padded
  // words (Hash::words zero-extends a partial last word), as the reader does
  void array(const void* p, size_t bytes) {
    put(p, bytes);
    static const unsigned char zero[8] = {0};
    if (bytes % 8) ok = ok && std::fwrite(zero, 1, 8 - bytes % 8, f) == 8 - bytes % 8;
  }
};

struct Reader {
  FILE* f;
  Hash hash;
  int64_t left;  // bytes of the file not read yet
  bool get(void* p, size_t bytes) {
    if (bytes == 0) return true;
    if (int64_t(bytes) > left || std::fread(p, 1, bytes, f) != bytes) return false;
    left -= int64_t(bytes);
    hash.words(p, bytes);
    return true;
  }
  template <typename T>
  bool array(T*& out, int64_t count) {
    const size_t bytes = size_t(count) * sizeof(T);
    const size_t padded = (bytes + 7) / 8 * 8;
    if (int64_t(padded) > left) return false;
    out = static_cast<T*>(std::malloc(padded ? padded : 8));
    if (!out) return false;
    if (!get(out, padded)) return false;
    return true;
  }
};

int32_t fail(const std::string& msg) {
  g_err = msg;
  return 4;
}

}  // namespace

extern "C" {

const char* emag_problem_last_error(void) { return g_err.c_str(); }

void emag_problem_free(emag_problem* p) {
  if (!p) return;
  std::free(p->xyz);
  std::free(p->tet2node);
  std::free(p->tet_material);
  std::free(p->tet_subdomain);
  std::free(p->nu);
  std::free(p->kappa);
  std::free(p->pec_edge_nodes);
  std::free(p->jq);
  std::memset(p, 0, sizeof(*p));
}

int32_t emag_problem_write(const char* path, const emag_problem* p) {
  g_err.clear();
  if (!path || !p) return fail("write: null argument");
  if (p->n_nodes <= 0 || p->n_tets <= 0 || p->n_pec < 0 || p->n_mat <= 0 || p->n_sub <= 0 ||
      !p->xyz || !p->tet2node || !p->tet_material || !p->tet_subdomain || !p->nu || !p->kappa ||
      (p->n_pec > 0 && !p->pec_edge_nodes))
    return fail("write: empty or incomplete problem");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("write: cannot open ") + path);
  Writer w{f};
  w.ok = std::fwrite(kMagic, 1, 8, f) == 8;
  const Header h{EMAG_PROBLEM_VERSION, p->jq ? 1u : 0u, p->n_nodes, p->n_tets, p->n_pec,
                 p->n_mat, p->n_sub};
  w.put(&h, sizeof h);
  w.array(p->xyz, sizeof(double) * 3 * size_t(p->n_nodes));
  w.array(p->tet2node, sizeof(int32_t) * 4 * size_t(p->n_tets));
  w.array(p->tet_material, sizeof(int32_t) * size_t(p->n_tets));
  w.array(p->tet_subdomain, sizeof(int32_t) * size_t(p->n_tets));
  w.array(p->nu, sizeof(double) * size_t(p->n_mat));
  w.array(p->kappa, sizeof(double) * size_t(p->n_mat));
  w.array(p->pec_edge_nodes, sizeof(int32_t) * 2 * size_t(p->n_pec));
  if (p->jq) w.array(p->jq, sizeof(double) * 12 * size_t(p->n_tets));
  const uint64_t sum = w.hash.h;
  w.ok = w.ok && std::fwrite(&sum, 1, 8, f) == 8;
  w.ok = (std::fclose(f) == 0) && w.ok;
  return w.ok ? 0 : fail(std::string("write: I/O error on ") + path);
}

int32_t emag_problem_read(const char* path, emag_problem* p) {
  g_err.clear();
  if (!path || !p) return fail("read: null argument");
  std::memset(p, 0, sizeof(*p));
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("read: cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const int64_t size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  char magic[8];
  Reader r{f, Hash{}, size};
  auto bad = [&](const std::string& msg) {
    std::fclose(f);
    emag_problem_free(p);
    return fail("read: " + msg + " (" + path + ")");
  };
  if (size < 8 + 40 + 8 || std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0)
    return bad("not an EMAGFDP problem file");
  r.left -= 8;
  Header h{};
  if (!r.get(&h, sizeof h)) return bad("truncated header");
  if (h.version != EMAG_PROBLEM_VERSION)
    return bad("unsupported version " + std::to_string(h.version));
  if (h.n_nodes <= 0 || h.n_tets <= 0 || h.n_pec < 0 || h.n_mat <= 0 || h.n_sub <= 0 ||
      h.n_nodes > INT32_MAX || h.n_tets > (int64_t(1) << 40) || h.n_pec > (int64_t(1) << 40))
    return bad("bad sizes in header");
  auto pad8 = [](int64_t b) { return (b + 7) / 8 * 8; };
  const int64_t need = pad8(24 * h.n_nodes) + pad8(16 * h.n_tets) + 2 * pad8(4 * h.n_tets) +
                       2 * pad8(8 * int64_t(h.n_mat)) + pad8(8 * h.n_pec) +
                       ((h.flags & 1) ? 96 * h.n_tets : 0) + 8;
  if (need != r.left) return bad("file length does not match its header");
  p->n_nodes = h.n_nodes;
  p->n_tets = h.n_tets;
  p->n_pec = h.n_pec;
  p->n_mat = h.n_mat;
  p->n_sub = h.n_sub;
  if (!r.array(p->xyz, 3 * h.n_nodes) || !r.array(p->tet2node, 4 * h.n_tets) ||
      !r.array(p->tet_material, h.n_tets) || !r.array(p->tet_subdomain, h.n_tets) ||
      !r.array(p->nu, h.n_mat) || !r.array(p->kappa, h.n_mat) ||
      !r.array(p->pec_edge_nodes, 2 * h.n_pec) ||
      ((h.flags & 1) && !r.array(p->jq, 12 * h.n_tets)))
    return bad("truncated arrays");
  uint64_t sum = 0;
  const uint64_t expect = r.hash.h;
  if (std::fread(&sum, 1, 8, f) != 8 || sum != expect) return bad("checksum mismatch");
  std::fclose(f);
  // index validation
  for (int64_t k = 0; k < 4 * h.n_tets; ++k)
    if (p->tet2node[k] < 0 || p->tet2node[k] >= h.n_nodes) {
      emag_problem_free(p);
      return fail("read: tet node index out of range");
    }
  std::vector<char> used(size_t(h.n_sub), 0);
  for (int64_t t = 0; t < h.n_tets; ++t) {
    if (p->tet_material[t] < 0 || p->tet_material[t] >= h.n_mat ||
        p->tet_subdomain[t] < 0 || p->tet_subdomain[t] >= h.n_sub) {
      emag_problem_free(p);
      return fail("read: material or subdomain id out of range");
    }
    used[size_t(p->tet_subdomain[t])] = 1;
  }
  for (int32_t s = 0; s < h.n_sub; ++s)
    if (!used[size_t(s)]) {
      emag_problem_free(p);
      return fail("read: subdomain " + std::to_string(s) + " has no tets");
    }
  for (int64_t k = 0; k < 2 * h.n_pec; ++k)
    if (p->pec_edge_nodes[k] < 0 || p->pec_edge_nodes[k] >= h.n_nodes) {
      emag_problem_free(p);
      return fail("read: PEC edge node out of range");
    }
  return 0;
}

}  // extern "C"
