// Mesh files: Wavefront OBJ input as the reference's tests read it, and a
// flattened file of a built mesh set (vertices, triangles, OBB trees) that
// loads without a tree build.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"

namespace mfcl {

namespace {

// Splits off the next token of s at any character of delims (empty tokens
// skipped); null when none is left.
char* nextToken(char*& s, const char* delims) {
  if (!s) return nullptr;
  s += std::strspn(s, delims);
  if (!*s) {
    s = nullptr;
    return nullptr;
  }
  char* tok = s;
  s += std::strcspn(s, delims);
  if (*s) *s++ = '\0';
  else s = nullptr;
  return tok;
}

}  // namespace

// test/test_fcl_utility.h:250-316 (loadOBJFile): lines of at most 2000
// bytes; a first token starting with '#' is a comment; "vn" and "vt" lines
// only set a flag; any other token starting with 'v' is a vertex (three
// coordinates); a token starting with 'f' is a face of k >= 3 ids, each the
// integer prefix of its token (atoi: "7/2/5" is 7), 1-based.  A face of k
// ids gives k - 2 triangles: once a "vn" or "vt" line has been read, the fan
// (0, t + 1, t + 2); before that, k - 2 copies of (0, 1, 2), as the
// reference's loader does.
int loadObj(const char* path, std::vector<float>& v, std::vector<uint32_t>& t, std::string& err) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    err = std::string("cannot open ") + path;
    return MFCL_E_INVALID;
  }
  bool fan = false;
  char buf[2000];
  uint64_t line = 0;
  int rc = MFCL_OK;
  while (rc == MFCL_OK && std::fgets(buf, sizeof(buf), f)) {
    ++line;
    char* s = buf;
    const char* first = nextToken(s, "\r\n\t ");
    if (!first || first[0] == '#') continue;
    if (first[0] == 'v') {
      if (first[1] == 'n' || first[1] == 't') {
        fan = true;
        continue;
      }
      for (int k = 0; k < 3; ++k) {
        const char* c = nextToken(s, "\t ");
        if (!c) {
          err = "line " + std::to_string(line) + ": vertex with fewer than 3 coordinates";
          rc = MFCL_E_INVALID;
          break;
        }
        v.push_back(static_cast<float>(std::atof(c)));
      }
    } else if (first[0] == 'f') {
      std::vector<long> ids;
      while (const char* c = nextToken(s, "\t \r\n")) ids.push_back(std::atol(c) - 1);
      if (ids.size() < 3) {
        err = "line " + std::to_string(line) + ": face with fewer than 3 vertices";
        rc = MFCL_E_INVALID;
        break;
      }
      for (size_t k = 0; k + 2 < ids.size(); ++k) {
        const long a = ids[0], b = fan ? ids[k + 1] : ids[1], c = fan ? ids[k + 2] : ids[2];
        if (a < 0 || b < 0 || c < 0) {
          err = "line " + std::to_string(line) + ": face vertex id below 1";
          rc = MFCL_E_INVALID;
          break;
        }
        t.push_back(static_cast<uint32_t>(a));
        t.push_back(static_cast<uint32_t>(b));
        t.push_back(static_cast<uint32_t>(c));
      }
    }
  }
  std::fclose(f);
  if (rc == MFCL_OK) {
    const uint64_t nv = v.size() / 3;
    for (uint32_t id : t)
      if (id >= nv) {
        err = "face vertex id beyond the vertex count";
        return MFCL_E_INVALID;
      }
  }
  return rc;
}

// ---- flattened mesh-set file --------------------------------------------
// Little-endian: "MFCLBVH" + '\0', uint32 version (1), uint32 mesh count,
// uint64 vertex, triangle and node totals, per mesh uint32 vertex,
// triangle and node counts, then the vertices (FP32 xyz), triangles (3
// uint32 local ids), 64-byte nodes, node ranges (2 uint32: first leaf-order
// slot relative to the mesh, count) and leaf-order triangle ids, and last an
// FNV-1a 64-bit hash of every byte before it.
namespace {

constexpr char kMagic[8] = {'M', 'F', 'C', 'L', 'B', 'V', 'H', '\0'};
constexpr uint32_t kVersion = 1;
constexpr size_t kNodeBytes = 64;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  }
};

struct Writer {
  FILE* f;
  Fnv fnv;
  bool ok = true;
  void put(const void* p, size_t n) {
    if (!n) return;
    fnv.add(p, n);
    ok = ok && std::fwrite(p, 1, n, f) == n;
  }
};

struct Reader {
  FILE* f;
  Fnv fnv;
  bool ok = true;
  void get(void* p, size_t n) {
    if (!n || !ok) return;
    ok = std::fread(p, 1, n, f) == n;
    if (ok) fnv.add(p, n);
  }
};

}  // namespace

int writeMeshFile(const MeshSetHost& m, const char* path, std::string& err) {
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    err = std::string("cannot create ") + path;
    return MFCL_E_INVALID;
  }
  Writer w{f};
  const uint32_t n = static_cast<uint32_t>(m.node_base.size());
  const uint64_t tot[3] = {m.verts.size() / 3, m.tris.size() / 3, m.nodes.size() / kNodeBytes};
  w.put(kMagic, 8);
  w.put(&kVersion, 4);
  w.put(&n, 4);
  w.put(tot, sizeof(tot));
  w.put(m.vert_count.data(), 4 * static_cast<size_t>(n));
  w.put(m.tri_count.data(), 4 * static_cast<size_t>(n));
  w.put(m.node_count.data(), 4 * static_cast<size_t>(n));
  w.put(m.verts.data(), 4 * m.verts.size());
  w.put(m.tris.data(), 4 * m.tris.size());
  w.put(m.nodes.data(), m.nodes.size());
  w.put(m.node_range.data(), 4 * m.node_range.size());
  w.put(m.prim.data(), 4 * m.prim.size());
  const uint64_t hash = w.fnv.h;
  w.ok = w.ok && std::fwrite(&hash, 1, 8, f) == 8;
  const bool ok = (std::fclose(f) == 0) && w.ok;
  if (!ok) {
    err = std::string("write error on ") + path;
    return MFCL_E_INVALID;
  }
  return MFCL_OK;
}

int readMeshFile(const char* path, MeshSetHost& m, std::string& err) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    err = std::string("cannot open ") + path;
    return MFCL_E_INVALID;
  }
  Reader r{f};
  auto bad = [&](const std::string& why) {
    std::fclose(f);
    err = std::string(path) + ": " + why;
    return MFCL_E_INVALID;
  };
  char magic[8];
  uint32_t version = 0, n = 0;
  uint64_t tot[3] = {0, 0, 0};
  r.get(magic, 8);
  r.get(&version, 4);
  r.get(&n, 4);
  r.get(tot, sizeof(tot));
  if (!r.ok || std::memcmp(magic, kMagic, 8) != 0) return bad("not a mesh-set file");
  if (version != kVersion) return bad("unsupported version " + std::to_string(version));
  if (n == 0 || tot[0] > (1ull << 34) || tot[1] > 0x7FFFFFFFull || tot[2] > 0x7FFFFFFFull)
    return bad("implausible sizes");
  m = MeshSetHost();
  m.vert_count.resize(n);
  m.tri_count.resize(n);
  m.node_count.resize(n);
  r.get(m.vert_count.data(), 4 * static_cast<size_t>(n));
  r.get(m.tri_count.data(), 4 * static_cast<size_t>(n));
  r.get(m.node_count.data(), 4 * static_cast<size_t>(n));
  if (!r.ok) return bad("truncated header");
  m.vert_base.resize(n);
  m.tri_base.resize(n);
  m.node_base.resize(n);
  uint64_t sv = 0, st = 0, sn = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (m.tri_count[i] == 0 || m.vert_count[i] == 0 || m.node_count[i] != 2 * m.tri_count[i] - 1)
      return bad("mesh " + std::to_string(i) + " has inconsistent counts");
    m.vert_base[i] = sv;
    m.tri_base[i] = static_cast<uint32_t>(st);
    m.node_base[i] = static_cast<uint32_t>(sn);
    sv += m.vert_count[i];
    st += m.tri_count[i];
    sn += m.node_count[i];
  }
  if (sv != tot[0] || st != tot[1] || sn != tot[2]) return bad("totals do not match the per-mesh counts");
  m.verts.resize(3 * sv);
  m.tris.resize(3 * st);
  m.nodes.resize(kNodeBytes * sn);
  m.node_range.resize(2 * sn);
  m.prim.resize(st);
  r.get(m.verts.data(), 4 * m.verts.size());
  r.get(m.tris.data(), 4 * m.tris.size());
  r.get(m.nodes.data(), m.nodes.size());
  r.get(m.node_range.data(), 4 * m.node_range.size());
  r.get(m.prim.data(), 4 * m.prim.size());
  uint64_t hash = 0;
  const uint64_t expect = r.fnv.h;
  const bool tail = r.ok && std::fread(&hash, 1, 8, f) == 8;
  std::fclose(f);
  if (!r.ok || !tail) {
    err = std::string(path) + ": truncated";
    return MFCL_E_INVALID;
  }
  if (hash != expect) {
    err = std::string(path) + ": checksum mismatch";
    return MFCL_E_INVALID;
  }
  // structure: every index in range, so no kernel reads outside the set
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t nt = m.tri_count[i], nv = m.vert_count[i], nn = m.node_count[i];
    for (uint64_t k = 0; k < 3ull * nt; ++k)
      if (m.tris[3ull * m.tri_base[i] + k] >= nv) {
        err = std::string(path) + ": triangle vertex id out of range in mesh " + std::to_string(i);
        return MFCL_E_INVALID;
      }
    for (uint32_t k = 0; k < nt; ++k)
      if (m.prim[m.tri_base[i] + k] >= nt) {
        err = std::string(path) + ": leaf-order triangle id out of range in mesh " + std::to_string(i);
        return MFCL_E_INVALID;
      }
    for (uint32_t k = 0; k < nn; ++k) {
      const size_t g = static_cast<size_t>(m.node_base[i]) + k;
      int32_t child;
      std::memcpy(&child, m.nodes.data() + kNodeBytes * g + 60, 4);
      const uint32_t first = m.node_range[2 * g], count = m.node_range[2 * g + 1];
      const bool range_ok = count >= 1 && first <= nt && count <= nt - first;
      const bool child_ok = child >= 0 ? (static_cast<uint32_t>(child) + 1 < nn && static_cast<uint32_t>(child) > k)
                                       : (static_cast<uint32_t>(-(child + 1)) < nt && count == 1);
      if (!range_ok || !child_ok) {
        err = std::string(path) + ": malformed node " + std::to_string(k) + " in mesh " + std::to_string(i);
        return MFCL_E_INVALID;
      }
    }
  }
  return MFCL_OK;
}

}  // namespace mfcl
