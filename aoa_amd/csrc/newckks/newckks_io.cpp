// Serialisation of polynomials, ciphertexts and keys.
//
// poly_export / poly_import keep the reference's text format
// (aoa-ntt/src/arithmetic/poly_t.cpp:426-473): every residue of the base in the
// coefficient domain as decimal numbers, each followed by a comma.
// cipher_export / cipher_import (aoa-ntt/src/ckks/cipher_t.cpp:42-82) and
// CKKSContext::export_keys / load_keys (aoa-ntt/src/ckks/ckkscontext.cpp:50-100)
// build and read the same documents, {"ct": {"0", "1"}, "props": {"level",
// "scale", "slots"}} and {"sk": {"s"}, "pk": {"a", "b"}, "evk": {"a", "b"}},
// as `json` trees (include/newckks/tool/json.h) instead of RapidJSON
// documents.  A binary form stores residues as little-endian uint64 words
// with a small header, in whichever domain the polynomials are (no transform
// on export).
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/newckks/newckks.h"

static void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

std::string poly_export(Context* ctx, poly_t* p) { return poly_residue_to_string(ctx, p); }

// Residues of the text format ("r0,r1,...,", spaces allowed) into h; returns
// how many were read.  A value that does not fit 64 bits, or more values than
// h holds, is an error.
static size_t parse_residues(const std::string& s, std::vector<uint64_t>* h, const char* who) {
  const size_t n = h->size();
  size_t idx = 0;
  uint64_t v = 0;
  bool any = false;
  auto push = [&]() {
    if (idx >= n) throw std::runtime_error(std::string(who) + ": more residues than the base holds");
    (*h)[idx++] = v;
    v = 0;
    any = false;
  };
  for (char c : s) {
    if (c >= '0' && c <= '9') {
      const uint64_t d = (uint64_t)(c - '0');
      if (v > (UINT64_MAX - d) / 10) throw std::runtime_error(std::string(who) + ": residue out of range");
      v = v * 10 + d;
      any = true;
    } else if (c == ',') {
      push();
    } else if (c != ' ' && c != '\n') {
      throw std::runtime_error(std::string(who) + ": unexpected character");
    }
  }
  if (any) push();
  return idx;
}

poly_t* poly_import(Context* ctx, std::string s) {
  poly_t* p = new poly_t;
  poly_init(ctx, p);
  std::vector<uint64_t> h((size_t)CUDAManager::get_n_residues(p->base) * CUDAManager::N, 0);
  parse_residues(s, &h, "poly_import");
  poly_copy_to_device(ctx, p, h.data());
  return p;
}

// Polynomial p of base b from every residue of that base in the text format,
// coefficient domain.
static void import_full(Context* ctx, poly_t* p, poly_bases b, const std::string& s, const char* who) {
  std::vector<uint64_t> h((size_t)CUDAManager::get_n_residues(b) * CUDAManager::N, 0);
  if (parse_residues(s, &h, who) != h.size())
    throw std::runtime_error(std::string(who) + ": fewer residues than the base holds");
  if (p->state != NONINITIALIZED && p->base != b) poly_free(ctx, p);
  poly_init(ctx, p, b);
  poly_copy_to_device(ctx, p, h.data());
}

// ------------------------------------------------------------ json ---------

json& json::operator[](const char* key) {
  if (kind_ == Null) kind_ = Object;
  if (kind_ != Object) throw std::runtime_error(std::string("json: not an object (member \"") + key + "\")");
  for (auto& m : members_)
    if (m.first == key) return m.second;
  members_.emplace_back(key, json());
  return members_.back().second;
}

const json& json::operator[](const char* key) const {
  if (kind_ == Object)
    for (const auto& m : members_)
      if (m.first == key) return m.second;
  throw std::runtime_error(std::string("json: missing member \"") + key + "\"");
}

bool json::HasMember(const char* key) const {
  if (kind_ != Object) return false;
  for (const auto& m : members_)
    if (m.first == key) return true;
  return false;
}

const char* json::GetString() const {
  if (kind_ != String) throw std::runtime_error("json: not a string");
  return str_.c_str();
}

int64_t json::GetInt64() const {
  if (kind_ != Int) throw std::runtime_error("json: not an integer");
  return int_;
}

int json::GetInt() const {
  const int64_t v = GetInt64();
  if (v < INT32_MIN || v > INT32_MAX) throw std::runtime_error("json: integer does not fit an int");
  return (int)v;
}

static void dump_string(const std::string& s, std::string* out) {
  out->push_back('"');
  for (char c : s) {
    switch (c) {
      case '"': *out += "\\\""; break;
      case '\\': *out += "\\\\"; break;
      case '\n': *out += "\\n"; break;
      case '\t': *out += "\\t"; break;
      case '\r': *out += "\\r"; break;
      default: out->push_back(c);
    }
  }
  out->push_back('"');
}

std::string json::dump() const {
  switch (kind_) {
    case String: {
      std::string o;
      dump_string(str_, &o);
      return o;
    }
    case Int: return std::to_string(int_);
    case Object: {
      std::string o = "{";
      for (size_t i = 0; i < members_.size(); ++i) {
        if (i) o += ",";
        dump_string(members_[i].first, &o);
        o += ":" + members_[i].second.dump();
      }
      return o + "}";
    }
    default: return "null";
  }
}

namespace {
struct Parser {
  const std::string& t;
  size_t i = 0;
  explicit Parser(const std::string& s) : t(s) {}
  [[noreturn]] void fail(const char* what) {
    throw std::runtime_error(std::string("json::parse: ") + what + " at offset " + std::to_string(i));
  }
  void ws() {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i;
  }
  std::string str() {
    if (i >= t.size() || t[i] != '"') fail("expected a string");
    ++i;
    std::string o;
    while (i < t.size() && t[i] != '"') {
      char c = t[i++];
      if (c == '\\') {
        if (i >= t.size()) fail("unterminated escape");
        const char e = t[i++];
        switch (e) {
          case '"': c = '"'; break;
          case '\\': c = '\\'; break;
          case '/': c = '/'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'n': c = '\n'; break;
          case 'r': c = '\r'; break;
          case 't': c = '\t'; break;
          default: fail("unsupported escape");
        }
      }
      o.push_back(c);
    }
    if (i >= t.size()) fail("unterminated string");
    ++i;
    return o;
  }
  json value() {
    ws();
    if (i >= t.size()) fail("unexpected end");
    if (t[i] == '{') {
      ++i;
      json o = json::object();
      ws();
      if (i < t.size() && t[i] == '}') {
        ++i;
        return o;
      }
      for (;;) {
        ws();
        const std::string k = str();
        ws();
        if (i >= t.size() || t[i] != ':') fail("expected ':'");
        ++i;
        o[k.c_str()] = value();
        ws();
        if (i < t.size() && t[i] == ',') {
          ++i;
          continue;
        }
        if (i < t.size() && t[i] == '}') {
          ++i;
          return o;
        }
        fail("expected ',' or '}'");
      }
    }
    if (t[i] == '"') return json::string(str());
    if (t[i] == '-' || (t[i] >= '0' && t[i] <= '9')) {
      const bool neg = t[i] == '-';
      if (neg) ++i;
      if (i >= t.size() || t[i] < '0' || t[i] > '9') fail("expected a digit");
      uint64_t v = 0;
      while (i < t.size() && t[i] >= '0' && t[i] <= '9') {
        const uint64_t d = (uint64_t)(t[i++] - '0');
        if (v > ((uint64_t)INT64_MAX + (neg ? 1 : 0) - d) / 10) fail("integer out of range");
        v = v * 10 + d;
      }
      if (i < t.size() && (t[i] == '.' || t[i] == 'e' || t[i] == 'E')) fail("only integers are supported");
      return json::integer(neg ? (int64_t)(0 - v) : (int64_t)v);
    }
    fail("unsupported value");
  }
};
}  // namespace

json json::parse(const std::string& text) {
  Parser p(text);
  json v = p.value();
  p.ws();
  if (p.i != text.size()) p.fail("trailing characters");
  return v;
}

// ------------------------------------------------------------ ciphertexts --

json cipher_export(Context* ctx, cipher_t* ct) {  // cipher_t.cpp:42-68
  json j = json::object();
  j["ct"]["0"] = json::string(poly_export(ctx, &ct->c[0]));
  j["ct"]["1"] = json::string(poly_export(ctx, &ct->c[1]));
  j["props"]["level"] = json::integer(ct->level);
  j["props"]["scale"] = json::integer(ct->scale);
  j["props"]["slots"] = json::integer(ct->slots);
  return j;
}

cipher_t* cipher_import(Context* ctx, const json& k) {  // cipher_t.cpp:70-82
  cipher_t* ct = new cipher_t;
  cipher_init(ctx, ct);
  try {
    for (int y = 0; y < 2; ++y) import_full(ctx, &ct->c[y], QBase, k["ct"][y ? "1" : "0"].GetString(), "cipher_import");
    ct->level = k["props"]["level"].GetInt();
    ct->scale = k["props"]["scale"].GetInt64();
    ct->slots = k["props"]["slots"].GetInt();
  } catch (...) {
    cipher_free(ctx, ct);
    delete ct;
    throw;
  }
  return ct;
}

std::string cipher_export_json(Context* ctx, cipher_t* ct) { return cipher_export(ctx, ct).dump(); }

cipher_t* cipher_import_json(Context* ctx, const std::string& doc) { return cipher_import(ctx, json::parse(doc)); }

// ------------------------------------------------------------ keys ---------

// The keys of the document, in its member order: (group, member, poly).
static std::vector<std::pair<std::pair<const char*, const char*>, poly_t*>> key_polys(CKKSKeychain& k) {
  return {{{"sk", "s"}, &k.sk->s}, {{"pk", "a"}, &k.pk->a}, {{"pk", "b"}, &k.pk->b},
          {{"evk", "a"}, &k.evk->a}, {{"evk", "b"}, &k.evk->b}};
}

static void need_keys(CKKSKeychain& k, const char* who) {
  if (!k.sk || !k.pk || !k.evk)
    throw std::runtime_error(std::string(who) + ": the context has no secret, public or evaluation key");
}

static void make_keys(CKKSKeychain& k) {
  if (!k.sk) k.sk = new SecretKey;
  if (!k.pk) k.pk = new PublicKey;
  if (!k.evk) k.evk = new EvaluationKey;
}

json CKKSContext::export_keys() {  // ckkscontext.cpp:60-100
  need_keys(keys, "export_keys");
  json j = json::object();
  for (auto& kp : key_polys(keys)) j[kp.first.first][kp.first.second] = json::string(poly_export(this, kp.second));
  return j;
}

// The reference reads each key string with poly_import, which makes a Q-base
// polynomial, so only the Q residues of its QB keys come back (and its host
// buffer is overrun by the special residues).  Here every key is read in
// full into a QB polynomial and transformed to the NTT domain, where the
// evaluation calls expect keys.
void CKKSContext::load_keys(const json& k) {  // ckkscontext.cpp:50-58
  make_keys(keys);
  for (auto& kp : key_polys(keys)) {
    import_full(this, kp.second, QBBase, k[kp.first.first][kp.first.second].GetString(), "load_keys");
    COMMONEngine::execute(this, kp.second, FORWARD);
  }
}

void CKKSContext::clear_keys() {  // ckkscontext.cpp:41-48
  need_keys(keys, "clear_keys");
  for (auto& kp : key_polys(keys)) poly_clear(this, kp.second);
}

// ------------------------------------------------------------ binary -------

namespace {
const char kMagic[8] = {'A', 'O', 'A', 'C', 'T', '0', '0', '1'};
const char kKeyMagic[8] = {'A', 'O', 'A', 'K', 'Y', '0', '0', '1'};
struct KeyHeader {
  char magic[8];
  int32_t N, k, kl, nkeys;
  int32_t state[8];
};
struct Header {
  char magic[8];
  int32_t N, k, level, slots;
  int64_t scale;
  int32_t state0, state1;
};
}  // namespace

std::vector<uint8_t> cipher_export_binary(Context* ctx, cipher_t* ct) {
  for (int y = 0; y < 2; ++y)
    if (ct->c[y].state == NONINITIALIZED) throw std::runtime_error("cipher_export_binary: uninitialised ciphertext");
  Header h;
  std::memcpy(h.magic, kMagic, 8);
  h.N = CUDAManager::N;
  h.k = (int32_t)CUDAManager::RNSQPrimes.size();
  h.level = ct->level;
  h.slots = ct->slots;
  h.scale = ct->scale;
  h.state0 = ct->c[0].state;
  h.state1 = ct->c[1].state;
  const size_t pb = (size_t)h.k * h.N * sizeof(uint64_t);
  std::vector<uint8_t> out(sizeof(Header) + 2 * pb);
  std::memcpy(out.data(), &h, sizeof(h));
  for (int y = 0; y < 2; ++y)
    hip_ok(hipMemcpyAsync(out.data() + sizeof(Header) + y * pb, ct->c[y].d_coefs, pb, hipMemcpyDeviceToHost,
                          ctx->get_stream()),
           "hipMemcpyAsync");
  hip_ok(hipStreamSynchronize(ctx->get_stream()), "hipStreamSynchronize");
  return out;
}

cipher_t* cipher_import_binary(Context* ctx, const uint8_t* data, size_t size) {
  Header h;
  if (size < sizeof(h)) throw std::runtime_error("cipher_import_binary: truncated header");
  std::memcpy(&h, data, sizeof(h));
  if (std::memcmp(h.magic, kMagic, 8) != 0) throw std::runtime_error("cipher_import_binary: not a ciphertext");
  if (h.N != CUDAManager::N || h.k != (int32_t)CUDAManager::RNSQPrimes.size())
    throw std::runtime_error("cipher_import_binary: parameters differ from the initialised ones");
  const size_t pb = (size_t)h.k * h.N * sizeof(uint64_t);
  if (size != sizeof(h) + 2 * pb) throw std::runtime_error("cipher_import_binary: wrong size");
  for (int32_t st : {h.state0, h.state1})
    if (st != RNSSTATE && st != TRANSSTATE) throw std::runtime_error("cipher_import_binary: bad domain flag");
  cipher_t* ct = new cipher_t;
  cipher_init(ctx, ct);
  for (int y = 0; y < 2; ++y) {
    hip_ok(hipMemcpyAsync(ct->c[y].d_coefs, data + sizeof(Header) + y * pb, pb, hipMemcpyHostToDevice,
                          ctx->get_stream()),
           "hipMemcpyAsync");
    ct->c[y].state = (poly_states)(y ? h.state1 : h.state0);
  }
  hip_ok(hipStreamSynchronize(ctx->get_stream()), "hipStreamSynchronize");
  ct->level = h.level;
  ct->slots = h.slots;
  ct->scale = h.scale;
  return ct;
}

// Keys in binary: a header (magic, N, k, kl, the number of polynomials and
// the domain of each) and sk.s, pk.a, pk.b, evk.a, evk.b as QB residues,
// little-endian uint64, in their current domain.
std::vector<uint8_t> CKKSContext::export_keys_binary() {
  need_keys(keys, "export_keys_binary");
  auto polys = key_polys(keys);
  KeyHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kKeyMagic, 8);
  h.N = CUDAManager::N;
  h.k = (int32_t)CUDAManager::RNSQPrimes.size();
  h.kl = (int32_t)CUDAManager::RNSBPrimes.size();
  h.nkeys = (int32_t)polys.size();
  const size_t pb = poly_get_size(QBBase);
  std::vector<uint8_t> out(sizeof(h) + polys.size() * pb);
  for (size_t y = 0; y < polys.size(); ++y) {
    poly_t* p = polys[y].second;
    if (p->state == NONINITIALIZED || p->base != QBBase)
      throw std::runtime_error("export_keys_binary: keys must be initialised QB polynomials");
    h.state[y] = p->state;
    hip_ok(hipMemcpyAsync(out.data() + sizeof(h) + y * pb, p->d_coefs, pb, hipMemcpyDeviceToHost, get_stream()),
           "hipMemcpyAsync");
  }
  std::memcpy(out.data(), &h, sizeof(h));
  hip_ok(hipStreamSynchronize(get_stream()), "hipStreamSynchronize");
  return out;
}

void CKKSContext::load_keys_binary(const uint8_t* data, size_t size) {
  KeyHeader h;
  if (size < sizeof(h)) throw std::runtime_error("load_keys_binary: truncated header");
  std::memcpy(&h, data, sizeof(h));
  if (std::memcmp(h.magic, kKeyMagic, 8) != 0) throw std::runtime_error("load_keys_binary: not a key file");
  if (h.N != CUDAManager::N || h.k != (int32_t)CUDAManager::RNSQPrimes.size() ||
      h.kl != (int32_t)CUDAManager::RNSBPrimes.size() || h.nkeys != 5)
    throw std::runtime_error("load_keys_binary: parameters differ from the initialised ones");
  const size_t pb = poly_get_size(QBBase);
  if (size != sizeof(h) + 5 * pb) throw std::runtime_error("load_keys_binary: wrong size");
  for (int y = 0; y < 5; ++y)
    if (h.state[y] != RNSSTATE && h.state[y] != TRANSSTATE) throw std::runtime_error("load_keys_binary: bad domain flag");
  make_keys(keys);
  auto polys = key_polys(keys);
  for (int y = 0; y < 5; ++y) {
    poly_t* p = polys[y].second;
    if (p->state != NONINITIALIZED && p->base != QBBase) poly_free(this, p);
    poly_init(this, p, QBBase);
    hip_ok(hipMemcpyAsync(p->d_coefs, data + sizeof(h) + y * pb, pb, hipMemcpyHostToDevice, get_stream()),
           "hipMemcpyAsync");
    p->state = (poly_states)h.state[y];
    COMMONEngine::execute(this, p, FORWARD);
  }
  hip_ok(hipStreamSynchronize(get_stream()), "hipStreamSynchronize");
}
