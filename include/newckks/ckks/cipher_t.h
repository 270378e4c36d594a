// Ciphertext (aoa-ntt/include/newckks/ckks/cipher_t.h:17-27).
#ifndef NEWCKKS_CIPHER_T_H
#define NEWCKKS_CIPHER_T_H

#include <stdint.h>

#include <string>
#include <vector>

#include "../arithmetic/poly_t.h"
#include "../tool/json.h"

struct ciphertext {
  poly_t c[2];
  int level;
  int slots;
  int64_t scale;
};
typedef ciphertext cipher_t;

void cipher_init(Context* ctx, cipher_t* c, int slots = 1);
void cipher_clear(Context* ctx, cipher_t* ct);
void cipher_free(Context* ctx, cipher_t* c);
void cipher_copy(Context* ctx, cipher_t* b, cipher_t* a);
// The reference's document (cipher_t.cpp:42-82):
// {"ct": {"0": "<poly_export>", "1": "..."}, "props": {"level", "scale", "slots"}}.
json cipher_export(Context* ctx, cipher_t* ct);
cipher_t* cipher_import(Context* ctx, const json& k);
// The same as compact text (json::dump / json::parse).
std::string cipher_export_json(Context* ctx, cipher_t* ct);
cipher_t* cipher_import_json(Context* ctx, const std::string& doc);
// Binary form (no counterpart in the reference): a 40-byte header (magic,
// N, k, level, slots, scale, domain of each polynomial) and the k limbs of
// both polynomials as little-endian uint64, in their current domain.
std::vector<uint8_t> cipher_export_binary(Context* ctx, cipher_t* ct);
cipher_t* cipher_import_binary(Context* ctx, const uint8_t* data, size_t size);

#endif
