// The document type of the serialisation API.  The reference passes
// RapidJSON documents (`typedef Document json;`,
// aoa-ntt/include/newckks/ckks/cipher_t.h:8-12); this layer has no RapidJSON,
// so `json` is a small ordered tree of objects, strings and integers with the
// part of the Document interface the reference's callers use on these
// documents (`k["ct"]["0"].GetString()`, `GetInt()`, `GetInt64()`,
// `HasMember`), plus `dump()` (the compact text RapidJSON's Writer produces:
// no whitespace, members in insertion order) and `json::parse`.
#ifndef NEWCKKS_JSON_H
#define NEWCKKS_JSON_H

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

class json {
 public:
  enum Kind { Null, Object, String, Int };

  json() = default;
  static json object() {
    json j;
    j.kind_ = Object;
    return j;
  }
  static json string(std::string s) {
    json j;
    j.kind_ = String;
    j.str_ = std::move(s);
    return j;
  }
  static json integer(int64_t v) {
    json j;
    j.kind_ = Int;
    j.int_ = v;
    return j;
  }

  // Member access; the non-const form adds a missing member (the object
  // keeps insertion order), the const form throws std::runtime_error.
  json& operator[](const char* key);
  const json& operator[](const char* key) const;
  bool HasMember(const char* key) const;

  bool IsObject() const { return kind_ == Object; }
  bool IsString() const { return kind_ == String; }
  bool IsInt64() const { return kind_ == Int; }
  // Throw std::runtime_error when the value has another type.
  const char* GetString() const;
  int GetInt() const;
  int64_t GetInt64() const;

  // Compact text: {"k":"v","n":1}.
  std::string dump() const;
  // Objects, strings (with \" \\ \/ \b \f \n \r \t escapes) and integers,
  // with any whitespace between tokens; throws std::runtime_error otherwise.
  static json parse(const std::string& text);

 private:
  Kind kind_ = Null;
  std::string str_;
  int64_t int_ = 0;
  std::vector<std::pair<std::string, json>> members_;
};

#endif
