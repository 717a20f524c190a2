// snarkjs .zkey / .wtns decoding and encoding (include/g16zkey.h). Host code.
#include <array>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/g16zkey.h"
#include "bn254.cuh"
#include "internal.h"

using namespace g16;

namespace {

constexpr uint32_t PROTOCOL_GROTH16 = 1;
constexpr size_t G1B = 64, G2B = 128, COEF_RECORD = 12 + 32;

struct Fail {
  std::string msg;
};

void le_modulus(uint8_t out[32], bool scalar) {
  for (int i = 0; i < 8; ++i) {
    uint32_t w = scalar ? FrCfg::p(i) : FqCfg::p(i);
    std::memcpy(out + 4 * i, &w, 4);
  }
}

// The whole file in memory plus its section index (ids may repeat or come in any order, so
// the chain is scanned once; crates/formats/src/binfile.rs:185-260).
struct BinFile {
  std::vector<uint8_t> data;
  std::vector<std::pair<uint32_t, std::pair<size_t, size_t>>> sections;  // id -> (off, len)

  void open(const char* path, const char magic[4], uint32_t max_version) {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw Fail{std::string("cannot open ") + path};
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = data.empty() ? 0 : std::fread(data.data(), 1, data.size(), f);
    std::fclose(f);
    if (got != data.size()) throw Fail{"short read"};
    if (data.size() < 12 || std::memcmp(data.data(), magic, 4) != 0)
      throw Fail{std::string("not a ") + std::string(magic, 4) + " file"};
    uint32_t version = u32(4), n = u32(8);
    if (version > max_version)
      throw Fail{"version " + std::to_string(version) + " exceeds supported " +
                 std::to_string(max_version)};
    size_t off = 12;
    for (uint32_t k = 0; k < n; ++k) {
      if (off + 12 > data.size()) throw Fail{"truncated section header"};
      uint32_t id = u32(off);
      uint64_t len;
      std::memcpy(&len, data.data() + off + 4, 8);
      off += 12;
      if (len > data.size() - off) throw Fail{"section " + std::to_string(id) + " overruns the file"};
      sections.push_back({id, {off, (size_t)len}});
      off += len;
    }
  }
  uint32_t u32(size_t off) const {
    uint32_t x;
    std::memcpy(&x, data.data() + off, 4);
    return x;
  }
  std::pair<size_t, size_t> unique(uint32_t id) const {
    int found = 0;
    std::pair<size_t, size_t> r{0, 0};
    for (auto& s : sections)
      if (s.first == id) {
        ++found;
        r = s.second;
      }
    if (found != 1)
      throw Fail{"section " + std::to_string(id) + (found ? " appears more than once" : " is missing")};
    return r;
  }
};

struct Cursor {
  const uint8_t* p;
  size_t left;
  uint32_t section;
  const uint8_t* take(size_t n) {
    if (n > left) throw Fail{"section " + std::to_string(section) + " is too short"};
    const uint8_t* r = p;
    p += n;
    left -= n;
    return r;
  }
  uint32_t u32() {
    uint32_t x;
    std::memcpy(&x, take(4), 4);
    return x;
  }
};

bool below(const uint8_t* le, bool scalar) {
  for (int i = 7; i >= 0; --i) {
    uint32_t w;
    std::memcpy(&w, le + 4 * i, 4);
    uint32_t m = scalar ? FrCfg::p(i) : FqCfg::p(i);
    if (w != m) return w < m;
  }
  return false;
}

void check_coords(const uint8_t* p, size_t n_fq, uint32_t section) {
  for (size_t i = 0; i < n_fq; ++i)
    if (!below(p + 32 * i, false))
      throw Fail{"section " + std::to_string(section) + ": a coordinate is not below the base field modulus"};
}

void check_modulus(Cursor& c, bool scalar) {
  uint32_t n8 = c.u32();
  uint8_t want[32];
  le_modulus(want, scalar);
  if (n8 != 32 || std::memcmp(c.take(32), want, 32) != 0)
    throw Fail{"unsupported curve (not BN254)"};
}

struct Hash32 {
  size_t operator()(const std::array<uint8_t, 32>& a) const {
    uint64_t h = 1469598103934665603ULL;
    for (uint8_t b : a) h = (h ^ b) * 1099511628211ULL;
    return (size_t)h;
  }
};

}  // namespace

struct g16zkey {
  g16hip_key key{};
  std::vector<uint8_t> header;  // alpha1 | beta1 | beta2 | gamma2 | delta1 | delta2
  std::vector<uint8_t> ic, values, a, b1, b2, l, h;
  std::vector<uint32_t> row_ptr[2], signal[2], value_id[2];
};

static void parse_zkey(const char* path, g16zkey* Z) {
  BinFile f;
  f.open(path, "zkey", 1);
  auto sec = [&](uint32_t id) {
    auto r = f.unique(id);
    return Cursor{f.data.data() + r.first, r.second, id};
  };
  Cursor s1 = sec(1);
  uint32_t protocol = s1.u32();
  if (protocol != PROTOCOL_GROTH16) throw Fail{"protocol " + std::to_string(protocol) + " is not groth16"};
  Cursor s2 = sec(2);
  check_modulus(s2, false);
  check_modulus(s2, true);
  uint32_t n_vars = s2.u32(), n_public = s2.u32(), domain = s2.u32();
  Z->header.assign(s2.take(3 * G1B + 3 * G2B), s2.p);
  if (s2.left) throw Fail{"trailing bytes after the groth16 header"};
  if (domain == 0 || (domain & (domain - 1))) throw Fail{"domain size " + std::to_string(domain) + " is not a power of two"};
  if (domain > (1u << 27)) throw Fail{"domain size exceeds 2^27"};
  if ((uint64_t)n_public + 1 > n_vars) throw Fail{"n_vars does not cover 1 + n_public"};
  check_coords(Z->header.data(), 2 * 3 + 4 * 3, 2);
  const uint8_t* hp = Z->header.data();
  const size_t hoff[6] = {0, 64, 128, 256, 384, 448};
  const size_t hlen[6] = {64, 64, 128, 128, 64, 128};
  for (int k = 0; k < 6; ++k) {
    bool zero = true;
    for (size_t i = 0; i < hlen[k]; ++i) zero &= hp[hoff[k] + i] == 0;
    if (zero) throw Fail{"a header point is the point at infinity, which no honest setup produces"};
  }
  auto points = [&](uint32_t id, size_t n, size_t width, std::vector<uint8_t>& out) {
    Cursor c = sec(id);
    if (c.left != n * width)
      throw Fail{"section " + std::to_string(id) + " holds " + std::to_string(c.left) +
                 " bytes, expected " + std::to_string(n * width)};
    out.assign(c.p, c.p + c.left);
    check_coords(out.data(), out.size() / 32, id);
  };
  points(3, (size_t)n_public + 1, G1B, Z->ic);
  points(5, n_vars, G1B, Z->a);
  points(6, n_vars, G1B, Z->b1);
  points(7, n_vars, G2B, Z->b2);
  points(8, (size_t)n_vars - n_public - 1, G1B, Z->l);
  points(9, domain, G1B, Z->h);

  // Section 4: (matrix, constraint, signal, v R^2) records -> CSR by constraint, file order
  // kept within a row, values interned (crates/formats/src/lib.rs:905-1130).
  Cursor s4 = sec(4);
  uint32_t n_coefs = s4.u32();
  if (s4.left != (size_t)n_coefs * COEF_RECORD) throw Fail{"section 4 record count does not match its length"};
  const uint8_t* rec = s4.p;
  for (int m = 0; m < 2; ++m) Z->row_ptr[m].assign((size_t)domain + 1, 0);
  for (uint32_t i = 0; i < n_coefs; ++i) {
    uint32_t m, c, sgn;
    std::memcpy(&m, rec + i * COEF_RECORD, 4);
    std::memcpy(&c, rec + i * COEF_RECORD + 4, 4);
    std::memcpy(&sgn, rec + i * COEF_RECORD + 8, 4);
    if (m > 1) throw Fail{"section 4: matrix id " + std::to_string(m) + " (only A and B are stored)"};
    if (c >= domain) throw Fail{"section 4: constraint " + std::to_string(c) + " is outside the domain"};
    if (sgn >= n_vars) throw Fail{"section 4: signal " + std::to_string(sgn) + " is out of range"};
    if (!below(rec + i * COEF_RECORD + 12, true)) throw Fail{"section 4: value is not below the scalar modulus"};
    Z->row_ptr[m][(size_t)c + 1]++;
  }
  for (int m = 0; m < 2; ++m)
    for (size_t c = 0; c < domain; ++c) Z->row_ptr[m][c + 1] += Z->row_ptr[m][c];
  std::vector<uint32_t> cursor[2] = {Z->row_ptr[0], Z->row_ptr[1]};
  for (int m = 0; m < 2; ++m) {
    Z->signal[m].resize(Z->row_ptr[m][domain]);
    Z->value_id[m].resize(Z->row_ptr[m][domain]);
  }
  std::unordered_map<std::array<uint8_t, 32>, uint32_t, Hash32> ids;
  for (uint32_t i = 0; i < n_coefs; ++i) {
    const uint8_t* r = rec + i * COEF_RECORD;
    uint32_t m, c, sgn;
    std::memcpy(&m, r, 4);
    std::memcpy(&c, r + 4, 4);
    std::memcpy(&sgn, r + 8, 4);
    // Stored integer v R^2; one Montgomery reduction (x R^-1) leaves v R, the Montgomery
    // form of v (crates/formats/src/binfile.rs:488-503).
    Fr stored;
    std::memcpy(stored.v, r + 12, 32);
    Fr mont = to_std(stored);
    std::array<uint8_t, 32> key;
    std::memcpy(key.data(), mont.v, 32);
    auto it = ids.find(key);
    uint32_t vid;
    if (it == ids.end()) {
      vid = (uint32_t)(Z->values.size() / 32);
      ids.emplace(key, vid);
      Z->values.insert(Z->values.end(), key.begin(), key.end());
    } else {
      vid = it->second;
    }
    uint32_t pos = cursor[m][c]++;
    Z->signal[m][pos] = sgn;
    Z->value_id[m][pos] = vid;
  }

  g16hip_key& k = Z->key;
  k.n_vars = n_vars;
  k.n_public = n_public;
  k.domain_size = domain;
  k.n_values = (uint32_t)(Z->values.size() / 32);
  k.alpha_g1 = hp + hoff[0];
  k.beta_g1 = hp + hoff[1];
  k.beta_g2 = hp + hoff[2];
  k.gamma_g2 = hp + hoff[3];
  k.delta_g1 = hp + hoff[4];
  k.delta_g2 = hp + hoff[5];
  k.ic = Z->ic.data();
  for (int m = 0; m < 2; ++m) {
    k.row_ptr[m] = Z->row_ptr[m].data();
    k.signal[m] = Z->signal[m].data();
    k.value_id[m] = Z->value_id[m].data();
  }
  k.values = Z->values.data();
  k.a_query = Z->a.data();
  k.b_g1_query = Z->b1.data();
  k.b_g2_query = Z->b2.data();
  k.l_query = Z->l.data();
  k.h_query = Z->h.data();
}

namespace {

struct Writer {
  std::vector<uint8_t> out;
  void raw(const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    out.insert(out.end(), b, b + n);
  }
  void u32(uint32_t x) { raw(&x, 4); }
  void u64(uint64_t x) { raw(&x, 8); }
  size_t begin(uint32_t id) {
    u32(id);
    size_t at = out.size();
    u64(0);
    return at;
  }
  void end(size_t at) {
    uint64_t len = out.size() - at - 8;
    std::memcpy(out.data() + at, &len, 8);
  }
  void save(const char* path) {
    FILE* f = std::fopen(path, "wb");
    if (!f) throw Fail{std::string("cannot create ") + path};
    size_t put = std::fwrite(out.data(), 1, out.size(), f);
    int rc = std::fclose(f);
    if (put != out.size() || rc != 0) throw Fail{std::string("short write to ") + path};
  }
};

}  // namespace

static void write_zkey(const char* path, const g16hip_key* k) {
  if (!k) throw Fail{"null key"};
  uint32_t n = k->domain_size;
  Writer w;
  w.raw("zkey", 4);
  w.u32(1);
  w.u32(10);
  size_t s = w.begin(1);
  w.u32(PROTOCOL_GROTH16);
  w.end(s);
  s = w.begin(2);
  uint8_t mod[32];
  w.u32(32);
  le_modulus(mod, false);
  w.raw(mod, 32);
  w.u32(32);
  le_modulus(mod, true);
  w.raw(mod, 32);
  w.u32(k->n_vars);
  w.u32(k->n_public);
  w.u32(n);
  w.raw(k->alpha_g1, G1B);
  w.raw(k->beta_g1, G1B);
  w.raw(k->beta_g2, G2B);
  w.raw(k->gamma_g2, G2B);
  w.raw(k->delta_g1, G1B);
  w.raw(k->delta_g2, G2B);
  w.end(s);
  s = w.begin(4);
  uint32_t total = k->row_ptr[0][n] + k->row_ptr[1][n];
  w.u32(total);
  for (int m = 0; m < 2; ++m)
    for (uint32_t c = 0; c < n; ++c)
      for (uint32_t e = k->row_ptr[m][c]; e < k->row_ptr[m][c + 1]; ++e) {
        w.u32((uint32_t)m);
        w.u32(c);
        w.u32(k->signal[m][e]);
        Fr mont;
        std::memcpy(mont.v, k->values + 32 * (size_t)k->value_id[m][e], 32);
        Fr stored = from_std(mont);  // v R -> v R^2, the section 4 encoding
        w.raw(stored.v, 32);
      }
  w.end(s);
  s = w.begin(3);
  w.raw(k->ic, G1B * ((size_t)k->n_public + 1));
  w.end(s);
  s = w.begin(9);
  w.raw(k->h_query, G1B * (size_t)n);
  w.end(s);
  s = w.begin(8);
  w.raw(k->l_query, G1B * ((size_t)k->n_vars - k->n_public - 1));
  w.end(s);
  s = w.begin(5);
  w.raw(k->a_query, G1B * (size_t)k->n_vars);
  w.end(s);
  s = w.begin(6);
  w.raw(k->b_g1_query, G1B * (size_t)k->n_vars);
  w.end(s);
  s = w.begin(7);
  w.raw(k->b_g2_query, G2B * (size_t)k->n_vars);
  w.end(s);
  s = w.begin(10);  // csHash (64 bytes) and no contributions
  uint8_t hash[64] = {0};
  w.raw(hash, 64);
  w.u32(0);
  w.end(s);
  w.save(path);
}

template <class F>
static g16hip_status catching(F&& f) {
  try {
    set_last_error("");
    f();
    return G16HIP_OK;
  } catch (const Fail& e) {
    set_last_error(e.msg);
    return G16HIP_BACKEND;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return G16HIP_BACKEND;
  }
}

extern "C" {

g16hip_status g16zkey_load(const char* path, g16zkey** out) {
  return catching([&] {
    if (!path || !out) throw Fail{"null argument"};
    auto Z = std::make_unique<g16zkey>();
    parse_zkey(path, Z.get());
    *out = Z.release();
  });
}

const g16hip_key* g16zkey_key(const g16zkey* z) { return z ? &z->key : nullptr; }

void g16zkey_free(g16zkey* z) { delete z; }

g16hip_status g16zkey_write(const char* path, const g16hip_key* key) {
  return catching([&] {
    if (!path) throw Fail{"null path"};
    write_zkey(path, key);
  });
}

g16hip_status g16wtns_load(const char* path, uint8_t** out, size_t* n_out) {
  return catching([&] {
    if (!path || !out || !n_out) throw Fail{"null argument"};
    BinFile f;
    f.open(path, "wtns", 2);
    auto r1 = f.unique(1);
    Cursor h{f.data.data() + r1.first, r1.second, 1};
    check_modulus(h, true);
    uint32_t n = h.u32();
    auto r2 = f.unique(2);
    if (r2.second != (size_t)n * 32) throw Fail{"section 2 does not hold n_witness values"};
    const uint8_t* src = f.data.data() + r2.first;
    std::unique_ptr<uint8_t[]> w(new uint8_t[(size_t)n * 32 + 1]);
    for (uint32_t i = 0; i < n; ++i) {
      if (!below(src + 32 * (size_t)i, true)) throw Fail{"witness value is not below the scalar modulus"};
      Fr plain;
      std::memcpy(plain.v, src + 32 * (size_t)i, 32);
      Fr mont = from_std(plain);
      std::memcpy(w.get() + 32 * (size_t)i, mont.v, 32);
    }
    if (n == 0 || !is_one(*reinterpret_cast<const Fr*>(w.get())))
      throw Fail{"witness[0] is not 1"};
    *out = w.release();
    *n_out = n;
  });
}

void g16wtns_free(uint8_t* w) { delete[] w; }

g16hip_status g16wtns_write(const char* path, const uint8_t* witness, size_t n) {
  return catching([&] {
    if (!path || (!witness && n)) throw Fail{"null argument"};
    Writer w;
    w.raw("wtns", 4);
    w.u32(2);
    w.u32(2);
    size_t s = w.begin(1);
    uint8_t mod[32];
    le_modulus(mod, true);
    w.u32(32);
    w.raw(mod, 32);
    w.u32((uint32_t)n);
    w.end(s);
    s = w.begin(2);
    for (size_t i = 0; i < n; ++i) {
      Fr mont;
      std::memcpy(mont.v, witness + 32 * i, 32);
      Fr plain = to_std(mont);
      w.raw(plain.v, 32);
    }
    w.end(s);
    w.save(path);
  });
}

}  // extern "C"
