// Serialisation of polynomials and ciphertexts.
//
// poly_export / poly_import keep the reference's text format
// (aoa-ntt/src/arithmetic/poly_t.cpp:445-473): every residue of the base in the
// coefficient domain as decimal numbers, each followed by a comma.  The
// reference's cipher_export / cipher_import (aoa-ntt/src/ckks/cipher_t.cpp:42-82)
// return a RapidJSON document; here the same JSON layout
// {"ct": {"0": "...", "1": "..."}, "props": {"level", "scale", "slots"}} is
// produced and parsed as a string.  A binary form stores the residues as
// little-endian uint64 words with a small header, in whichever domain the
// polynomials are (no transform on export).
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

poly_t* poly_import(Context* ctx, std::string s) {
  poly_t* p = new poly_t;
  poly_init(ctx, p);
  const size_t n = (size_t)CUDAManager::get_n_residues(p->base) * CUDAManager::N;
  std::vector<uint64_t> h(n, 0);
  size_t idx = 0;
  uint64_t v = 0;
  bool any = false;
  for (char c : s) {
    if (c >= '0' && c <= '9') {
      v = v * 10 + (uint64_t)(c - '0');
      any = true;
    } else if (c == ',') {
      if (idx >= n) throw std::runtime_error("poly_import: more residues than the base holds");
      h[idx++] = v;
      v = 0;
      any = false;
    } else if (c != ' ' && c != '\n') {
      throw std::runtime_error("poly_import: unexpected character");
    }
  }
  if (any) {
    if (idx >= n) throw std::runtime_error("poly_import: more residues than the base holds");
    h[idx++] = v;
  }
  poly_copy_to_device(ctx, p, h.data());
  return p;
}

// ------------------------------------------------------------ JSON ---------

static std::string json_field(const std::string& s, const std::string& key, size_t from = 0) {
  const std::string k = "\"" + key + "\"";
  size_t i = s.find(k, from);
  if (i == std::string::npos) throw std::runtime_error("cipher_import_json: missing \"" + key + "\"");
  i = s.find(':', i + k.size());
  if (i == std::string::npos) throw std::runtime_error("cipher_import_json: malformed document");
  ++i;
  while (i < s.size() && (s[i] == ' ' || s[i] == '\n')) ++i;
  if (i < s.size() && s[i] == '"') {
    const size_t e = s.find('"', i + 1);
    if (e == std::string::npos) throw std::runtime_error("cipher_import_json: unterminated string");
    return s.substr(i + 1, e - i - 1);
  }
  size_t e = i;
  while (e < s.size() && (s[e] == '-' || (s[e] >= '0' && s[e] <= '9'))) ++e;
  return s.substr(i, e - i);
}

std::string cipher_export_json(Context* ctx, cipher_t* ct) {
  return "{\"ct\":{\"0\":\"" + poly_export(ctx, &ct->c[0]) + "\",\"1\":\"" + poly_export(ctx, &ct->c[1]) +
         "\"},\"props\":{\"level\":" + std::to_string(ct->level) + ",\"scale\":" + std::to_string(ct->scale) +
         ",\"slots\":" + std::to_string(ct->slots) + "}}";
}

cipher_t* cipher_import_json(Context* ctx, const std::string& doc) {
  cipher_t* ct = new cipher_t;
  cipher_init(ctx, ct);
  const size_t c = doc.find("\"ct\"");
  if (c == std::string::npos) throw std::runtime_error("cipher_import_json: missing \"ct\"");
  for (int y = 0; y < 2; ++y) {
    poly_t* p = poly_import(ctx, json_field(doc, y ? "1" : "0", c));
    poly_copy(ctx, &ct->c[y], p);
    poly_free(ctx, p);
    delete p;
  }
  const size_t pr = doc.find("\"props\"");
  if (pr == std::string::npos) throw std::runtime_error("cipher_import_json: missing \"props\"");
  ct->level = std::stoi(json_field(doc, "level", pr));
  ct->scale = std::stoll(json_field(doc, "scale", pr));
  ct->slots = std::stoi(json_field(doc, "slots", pr));
  return ct;
}

// ------------------------------------------------------------ binary -------

namespace {
const char kMagic[8] = {'A', 'O', 'A', 'C', 'T', '0', '0', '1'};
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
