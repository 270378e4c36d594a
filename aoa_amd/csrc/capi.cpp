// C-ABI entry points (include/aoa.h): argument checks, workspace carving and
// the kernel sequences of the plain (one ciphertext per call) form.
#include <algorithm>
#include <cstring>
#include <string>

#include "internal.hpp"

using namespace aoa;

static thread_local std::string g_err;

namespace aoa {

int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(AOA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return AOA_OK;
}

uint64_t galois_of(int rot_slots, int N) {
  const uint64_t M = 2 * (uint64_t)N;
  uint64_t r = 1, s = 5 % M, e = (uint64_t)rot_slots;
  while (e) {
    if (e & 1) r = (r * s) % M;
    s = (s * s) % M;
    e >>= 1;
  }
  return r;
}

}  // namespace aoa

// Ciphertext and workspace pointers are moved as 16-byte vectors.
static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

static int fail(int code, const std::string& msg) { return set_error(code, msg); }

extern "C" const char* aoa_last_error(void) { return g_err.c_str(); }

extern "C" int aoa_params_create(int N, int k, int kl, int prec, aoa_params** out) {
  if (!out) return fail(AOA_ERR_ARG, "null output handle");
  *out = nullptr;
  aoa_params* h = new aoa_params();
  std::string err = build_host_params(N, k, kl, prec, &h->host);
  if (!err.empty()) {
    delete h;
    return fail(AOA_ERR_UNSUPPORTED, err);
  }
  const HostParams& H = h->host;
  // One allocation, 256-byte aligned sections.
  struct Sec {
    const void* src;
    size_t bytes;
    size_t off;
  };
  std::vector<Sec> secs = {
      {H.dprimes.data(), H.dprimes.size() * sizeof(DevPrime), 0},
      {H.tw_fwd.data(), H.tw_fwd.size() * 8, 0},
      {H.tw_inv.data(), H.tw_inv.size() * 8, 0},
      {H.inv_qhat.data(), H.inv_qhat.size() * 8, 0},
      {H.qhat_mod_b.data(), H.qhat_mod_b.size() * 8, 0},
      {H.inv_bhat.data(), H.inv_bhat.size() * 8, 0},
      {H.bhat_mod_q.data(), H.bhat_mod_q.size() * 8, 0},
      {H.inv_b_mod_q.data(), H.inv_b_mod_q.size() * 8, 0},
      {H.b_mod_q.data(), H.b_mod_q.size() * 8, 0},
      {H.inv_ql_mod_qi.data(), H.inv_ql_mod_qi.size() * 8, 0},
      {H.enc_w.data(), H.enc_w.size() * 8, 0},
      {H.rot_group.data(), H.rot_group.size() * 4, 0},
      {H.gauss_cdt.data(), H.gauss_cdt.size() * 8, 0},
      {H.md_terms.data(), H.md_terms.size() * sizeof(MdTerm), 0},
      {H.md_thr.data(), H.md_thr.size() * 4, 0},
      {H.md_red.data(), H.md_red.size() * 8, 0},
      {H.md_fin.data(), H.md_fin.size() * 8, 0},
      {H.md_negb.data(), H.md_negb.size() * 4, 0},
      {H.mu_terms.data(), H.mu_terms.size() * 4, 0},
      {H.bscale.data(), H.bscale.size() * 8, 0},
  };
  // Row-pass twiddle orders, one per two-pass split of N and direction
  // (device only: the host keeps the natural tables).
  std::vector<uint64_t> rowtw[3][2];
  int rowsec[3][2];
  for (int si = 0; si < 3; ++si)
    for (int dir = 0; dir < 2; ++dir) {
      row_twiddle_order(dir ? H.tw_inv : H.tw_fwd, H.N, H.k + H.K, 64 << si, &rowtw[si][dir]);
      rowsec[si][dir] = rowtw[si][dir].empty() ? -1 : (int)secs.size();
      if (!rowtw[si][dir].empty()) secs.push_back({rowtw[si][dir].data(), rowtw[si][dir].size() * 8, 0});
    }
  const int nsec = (int)secs.size();
  size_t total = 0;
  for (int i = 0; i < nsec; ++i) {
    secs[i].off = total;
    total += (secs[i].bytes + 255) & ~(size_t)255;
  }
  if (hipMalloc(&h->dmem, total) != hipSuccess) {
    delete h;
    return fail(AOA_ERR_HIP, "hipMalloc of parameter tables failed");
  }
  for (int i = 0; i < nsec; ++i) {
    if (hipMemcpy((char*)h->dmem + secs[i].off, secs[i].src, secs[i].bytes, hipMemcpyHostToDevice) !=
        hipSuccess) {
      (void)hipFree(h->dmem);
      delete h;
      return fail(AOA_ERR_HIP, "hipMemcpy of parameter tables failed");
    }
  }
  char* b = (char*)h->dmem;
  DevParams& D = h->dev;
  D.N = H.N;
  D.logN = H.logN;
  D.k = H.k;
  D.K = H.K;
  D.primes = (const DevPrime*)(b + secs[0].off);
  D.tw_fwd = (const uint64_t*)(b + secs[1].off);
  D.tw_inv = (const uint64_t*)(b + secs[2].off);
  D.inv_qhat = (const uint64_t*)(b + secs[3].off);
  D.qhat_mod_b = (const uint64_t*)(b + secs[4].off);
  D.inv_bhat = (const uint64_t*)(b + secs[5].off);
  D.bhat_mod_q = (const uint64_t*)(b + secs[6].off);
  D.inv_b_mod_q = (const uint64_t*)(b + secs[7].off);
  D.b_mod_q = (const uint64_t*)(b + secs[8].off);
  D.inv_ql_mod_qi = (const uint64_t*)(b + secs[9].off);
  D.enc_w = (const double*)(b + secs[10].off);
  D.rot_group = (const uint32_t*)(b + secs[11].off);
  D.gauss_cdt = (const uint64_t*)(b + secs[12].off);
  D.md_terms = (const uint32_t*)(b + secs[13].off);
  D.md_thr = (const uint32_t*)(b + secs[14].off);
  D.md_red = (const uint64_t*)(b + secs[15].off);
  D.md_fin = (const uint64_t*)(b + secs[16].off);
  D.md_negb = (const uint32_t*)(b + secs[17].off);
  D.mu_terms = (const uint32_t*)(b + secs[18].off);
  D.bscale = (const uint64_t*)(b + secs[19].off);
  for (int si = 0; si < 3; ++si)
    for (int dir = 0; dir < 2; ++dir)
      D.tw_row[si][dir] = rowsec[si][dir] < 0 ? nullptr : (const uint64_t*)(b + secs[rowsec[si][dir]].off);
  *out = h;
  return AOA_OK;
}

extern "C" void aoa_params_destroy(aoa_params* h) {
  if (!h) return;
  if (h->dmem) (void)hipFree(h->dmem);
  delete h;
}

extern "C" int aoa_params_info(const aoa_params* h, int* N, int* k, int* kl, int* prec) {
  if (!h) return fail(AOA_ERR_ARG, "null params");
  if (N) *N = h->host.N;
  if (k) *k = h->host.k;
  if (kl) *kl = h->host.K;
  if (prec) *prec = h->host.prec;
  return AOA_OK;
}

extern "C" int aoa_params_primes(const aoa_params* h, uint64_t* out, int n) {
  if (!h || !out) return fail(AOA_ERR_ARG, "null argument");
  int R = (int)h->host.primes.size();
  if (n < R) return fail(AOA_ERR_ARG, "output too small");
  std::memcpy(out, h->host.primes.data(), R * sizeof(uint64_t));
  return AOA_OK;
}

extern "C" int64_t aoa_params_table(const aoa_params* h, int which, uint64_t* out, int64_t n) {
  if (!h) return -fail(AOA_ERR_ARG, "null params");
  const HostParams& H = h->host;
  const int R = H.k + H.K;
  std::vector<uint64_t> v;
  auto take = [&](const std::vector<uint64_t>& src, int stride) {
    for (size_t i = 0; i < src.size(); i += stride) v.push_back(src[i]);
  };
  switch (which) {
    case 0: take(H.tw_fwd, 2); break;
    case 1: take(H.tw_inv, 2); break;
    case 2: take(H.inv_qhat, 2); break;
    case 3: take(H.qhat_mod_b, 1); break;
    case 4: take(H.inv_bhat, 2); break;
    case 5: take(H.bhat_mod_q, 2); break;
    case 6: take(H.inv_b_mod_q, 2); break;
    case 7: take(H.b_mod_q, 2); break;
    case 8: take(H.inv_ql_mod_qi, 2); break;
    case 9:
      for (int r = 0; r < R; ++r) v.push_back(H.dprimes[r].mu);
      break;
    case 11:
      v.resize(H.enc_w.size());
      std::memcpy(v.data(), H.enc_w.data(), v.size() * 8);
      break;
    case 12:
      for (uint32_t g : H.rot_group) v.push_back(g);
      break;
    case 13: v = H.gauss_cdt; break;
    case 15: take(H.bscale, 2); break;
    case 14:  // ModDown constants, eight 32-bit words per (i, j) widened to 64 bits
      for (const MdConst& m : H.md_const)
        for (uint32_t x : {m.cl, m.ch, m.cs, m.w0, m.w1, m.w2, m.thr, m.pad}) v.push_back(x);
      break;
    default: return -fail(AOA_ERR_ARG, "unknown table");
  }
  if (!out) return (int64_t)v.size();
  if (n < (int64_t)v.size()) return -fail(AOA_ERR_ARG, "output too small");
  std::memcpy(out, v.data(), v.size() * 8);
  return (int64_t)v.size();
}

static LimbMap limbs_of(uint64_t* base, int res0, int nres, int k) {
  LimbMap m;
  std::memset(&m, 0, sizeof(m));
  m.base[0] = base + 0;
  m.nlimbs = nres;
  m.npolys = 1;
  if (res0 < k) {
    m.prime0 = res0;
    m.qlimbs = std::min(nres, k - res0);
  } else {
    // only special limbs: prime = k + li - qlimbs with qlimbs = k - res0 <= 0
    m.prime0 = 0;
    m.qlimbs = k - res0;
  }
  return m;
}

extern "C" int aoa_ntt(const aoa_params* h, void* stream, uint64_t* d, int res0, int nres, int dir) {
  if (!h || !d) return fail(AOA_ERR_ARG, "null argument");
  const int R = h->host.k + h->host.K;
  if (res0 < 0 || nres <= 0 || res0 + nres > R) return fail(AOA_ERR_ARG, "limb range out of bounds");
  LimbMap m = limbs_of(d + (size_t)res0 * h->host.N, res0, nres, h->host.k);
  hipStream_t st = (hipStream_t)stream;
  int rc = dir == AOA_FORWARD ? ntt_forward(h->dev, m, 0, st) : ntt_inverse(h->dev, m, 0, nullptr, nullptr, st);
  if (rc) return fail(AOA_ERR_UNSUPPORTED, "no NTT split for this N");
  return check_launch("aoa_ntt");
}

extern "C" int aoa_eltwise(const aoa_params* h, void* stream, int op, uint64_t* out, const uint64_t* a,
                           const uint64_t* b, const uint64_t* c, int res0, int nres) {
  if (!h || !out || !a) return fail(AOA_ERR_ARG, "null argument");
  if (op != AOA_ADD && op != AOA_SUB && op != AOA_MUL && op != AOA_NEGATE && op != AOA_MULADD)
    return fail(AOA_ERR_ARG, "unknown operation");
  if (op != AOA_NEGATE && !b) return fail(AOA_ERR_ARG, "null operand b");
  if (op == AOA_MULADD && !c) return fail(AOA_ERR_ARG, "null addend c");
  const int R = h->host.k + h->host.K, k = h->host.k;
  if (res0 < 0 || nres <= 0 || res0 + nres > R) return fail(AOA_ERR_ARG, "limb range out of bounds");
  const size_t off = (size_t)res0 * h->host.N;
  int qlimbs = res0 < k ? std::min(nres, k - res0) : k - res0;
  launch_eltwise(h->dev, op, out + off, a + off, b ? b + off : nullptr, c ? c + off : nullptr, nres, qlimbs,
                 res0 < k ? res0 : 0, (hipStream_t)stream);
  return check_launch("aoa_eltwise");
}

extern "C" int aoa_eltwise_scalar(const aoa_params* h, void* stream, int op, uint64_t* out, const uint64_t* a,
                                  uint64_t b, int res0, int nres) {
  if (!h || !out || !a) return fail(AOA_ERR_ARG, "null argument");
  if (op != AOA_ADD && op != AOA_SUB && op != AOA_MUL) return fail(AOA_ERR_ARG, "unknown operation");
  const int R = h->host.k + h->host.K;
  if (res0 < 0 || nres <= 0 || res0 + nres > R) return fail(AOA_ERR_ARG, "limb range out of bounds");
  const size_t off = (size_t)res0 * h->host.N;
  launch_eltwise_scalar(h->dev, op, out + off, a + off, b, nres, res0, (hipStream_t)stream);
  return check_launch("aoa_eltwise_scalar");
}

extern "C" int aoa_conjugate(const aoa_params* h, void* stream, uint64_t* out, const uint64_t* in, int nres) {
  if (!h || !out || !in) return fail(AOA_ERR_ARG, "null argument");
  if (out == in) return fail(AOA_ERR_ARG, "out and in must not overlap");
  if (nres <= 0 || nres > h->host.k + h->host.K) return fail(AOA_ERR_ARG, "limb count out of range");
  launch_conjugate(h->dev, out, in, nres, (hipStream_t)stream);
  return check_launch("aoa_conjugate");
}

static bool slots_ok(const aoa_params* h, int n) { return n >= 1 && n <= h->host.N / 2 && !(n & (n - 1)); }

extern "C" int aoa_special_fft(const aoa_params* h, void* stream, double* vals, int n, int dir) {
  if (!h || !vals) return fail(AOA_ERR_ARG, "null argument");
  if (!slots_ok(h, n)) return fail(AOA_ERR_ARG, "size must be a power of two <= N/2");
  if (dir != AOA_FORWARD && dir != AOA_INVERSE) return fail(AOA_ERR_ARG, "unknown direction");
  if (!aligned16(vals)) return fail(AOA_ERR_ARG, "values must be 16-byte aligned");
  launch_special_fft(h->dev, vals, n, dir == AOA_INVERSE, (hipStream_t)stream);
  return check_launch("aoa_special_fft");
}

extern "C" int aoa_encode(const aoa_params* h, void* stream, uint64_t* out, double* vals, int slots, uint64_t scale,
                          int nres) {
  if (!h || !out || !vals) return fail(AOA_ERR_ARG, "null argument");
  if (!slots_ok(h, slots)) return fail(AOA_ERR_ARG, "slots must be a power of two <= N/2");
  if (scale == 0) return fail(AOA_ERR_ARG, "scale must be positive");
  if (nres <= 0 || nres > h->host.k) return fail(AOA_ERR_ARG, "limb count out of range");
  if (!aligned16(vals)) return fail(AOA_ERR_ARG, "values must be 16-byte aligned");
  launch_special_fft(h->dev, vals, slots, 1, (hipStream_t)stream);
  launch_encode_pack(h->dev, out, vals, slots, (double)scale, nres, (hipStream_t)stream);
  return check_launch("aoa_encode");
}

extern "C" int aoa_decode(const aoa_params* h, void* stream, double* vals, const uint64_t* in, int slots,
                          uint64_t scale) {
  if (!h || !vals || !in) return fail(AOA_ERR_ARG, "null argument");
  if (!slots_ok(h, slots)) return fail(AOA_ERR_ARG, "slots must be a power of two <= N/2");
  if (scale == 0) return fail(AOA_ERR_ARG, "scale must be positive");
  if (!aligned16(vals)) return fail(AOA_ERR_ARG, "values must be 16-byte aligned");
  launch_decode_unpack(h->dev, vals, in, slots, (double)scale, (hipStream_t)stream);
  launch_special_fft(h->dev, vals, slots, 0, (hipStream_t)stream);
  return check_launch("aoa_decode");
}

extern "C" int aoa_sample(const aoa_params* h, void* stream, uint64_t* out, int res0, int nres, int kind,
                          uint64_t seed, uint64_t nonce) {
  if (!h || !out) return fail(AOA_ERR_ARG, "null argument");
  if (kind < AOA_SAMPLE_UNIFORM || kind > AOA_SAMPLE_NARROW) return fail(AOA_ERR_ARG, "unknown distribution");
  const int R = h->host.k + h->host.K;
  if (res0 < 0 || nres <= 0 || res0 + nres > R) return fail(AOA_ERR_ARG, "limb range out of bounds");
  launch_sample(h->dev, out, res0, nres, kind, seed, nonce, (hipStream_t)stream);
  return check_launch("aoa_sample");
}

extern "C" int aoa_encrypt(const aoa_params* h, void* stream, uint64_t* c0, uint64_t* c1, uint64_t* m,
                           const uint64_t* pk_b, const uint64_t* pk_a, uint64_t seed, uint64_t nonce, void* ws,
                           size_t ws_bytes) {
  if (!h || !c0 || !c1 || !m || !pk_b || !pk_a || !ws) return fail(AOA_ERR_ARG, "null argument");
  const int k = h->host.k;
  const size_t N = (size_t)h->host.N;
  if (ws_bytes < 3 * (size_t)k * N * sizeof(uint64_t)) return fail(AOA_ERR_WORKSPACE, "workspace too small");
  const DevParams& P = h->dev;
  hipStream_t st = (hipStream_t)stream;
  uint64_t* u = (uint64_t*)ws;
  uint64_t* e0 = u + (size_t)k * N;
  uint64_t* e1 = e0 + (size_t)k * N;
  launch_sample_encrypt(P, u, e0, e1, k, seed, nonce, st);
  LimbMap mm;
  std::memset(&mm, 0, sizeof(mm));
  mm.base[0] = m;
  mm.base[1] = u;
  mm.base[2] = e0;
  mm.base[3] = e1;
  mm.nlimbs = k;
  mm.qlimbs = k;
  mm.npolys = 4;
  if (ntt_forward(P, mm, 0, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split for this N");
  launch_encrypt_combine(P, c0, c1, pk_b, pk_a, u, e0, e1, m, k, st);
  return check_launch("aoa_encrypt");
}

extern "C" int aoa_decrypt(const aoa_params* h, void* stream, uint64_t* m, uint64_t* c0, uint64_t* c1,
                           const uint64_t* s, int level) {
  if (!h || !m || !c0 || !c1 || !s) return fail(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  const DevParams& P = h->dev;
  hipStream_t st = (hipStream_t)stream;
  const int L = level + 1;
  LimbMap mc;
  std::memset(&mc, 0, sizeof(mc));
  mc.base[0] = c0;
  mc.base[1] = c1;
  mc.nlimbs = L;
  mc.qlimbs = L;
  mc.npolys = 2;
  if (ntt_forward(P, mc, 0, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split for this N");
  launch_eltwise(P, ELT_MULADD, m, c1, s, c0, L, L, 0, st);
  LimbMap mo = limbs_of(m, 0, L, h->host.k);
  if (ntt_inverse(P, mo, 0, nullptr, nullptr, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split for this N");
  return check_launch("aoa_decrypt");
}

extern "C" int aoa_rescale(const aoa_params* h, void* stream, uint64_t* c0, uint64_t* c1, int level) {
  if (!h || !c0 || !c1) return fail(AOA_ERR_ARG, "null argument");
  if (level == 0) return fail(AOA_ERR_LEVEL, "The ciphertext level is 0. No more rescaling is possible.");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  launch_rescale(h->dev, c0, c1, level, (hipStream_t)stream);
  return check_launch("aoa_rescale");
}

extern "C" int aoa_modup(const aoa_params* h, void* stream, uint64_t* d_qb, int level) {
  if (!h || !d_qb) return fail(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  launch_modup(h->dev, d_qb, d_qb + (size_t)h->host.k * h->host.N, level, (hipStream_t)stream);
  return check_launch("aoa_modup");
}

extern "C" int aoa_modup_from(const aoa_params* h, void* stream, uint64_t* d_qb, const uint64_t* d_q, int level) {
  if (!h || !d_qb || !d_q) return fail(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  const size_t N = (size_t)h->host.N, k = (size_t)h->host.k;
  // the special limbs written must not overlap the Q limbs read (unless they
  // are the same QB polynomial)
  const uint64_t* w0 = d_qb + k * N;
  const uint64_t* w1 = w0 + (size_t)h->host.K * N;
  const uint64_t* r1 = d_q + (size_t)(level + 1) * N;
  if (d_q != d_qb && r1 > w0 && d_q < w1) return fail(AOA_ERR_ARG, "input overlaps the special limbs written");
  launch_modup(h->dev, d_q, d_qb + k * N, level, (hipStream_t)stream);
  return check_launch("aoa_modup_from");
}

extern "C" int aoa_dot(const aoa_params* h, void* stream, uint64_t* c, const uint64_t* const* a,
                       const uint64_t* const* b, int k) {
  if (!h || !c || !a || !b) return fail(AOA_ERR_ARG, "null argument");
  if (k < 0 || k > 1 << 20) return fail(AOA_ERR_ARG, "term count out of range");
  launch_dot(h->dev, c, a, b, k, (hipStream_t)stream);
  return check_launch("aoa_dot");
}

extern "C" int aoa_dr2(const aoa_params* h, void* stream, uint64_t* c21, uint64_t* c22, uint64_t* c23,
                       const uint64_t* a01, const uint64_t* a02, const uint64_t* a11, const uint64_t* a12) {
  if (!h || !c21 || !c22 || !c23 || !a01 || !a02 || !a11 || !a12) return fail(AOA_ERR_ARG, "null argument");
  launch_dr2(h->dev, c21, c22, c23, a01, a02, a11, a12, (hipStream_t)stream);
  return check_launch("aoa_dr2");
}

extern "C" int aoa_rns_primes(int k, int kl, int prec, uint64_t* out, int n) {
  if (kl <= 0) kl = k + 1;
  std::vector<uint64_t> pr;
  const std::string err = select_rns_primes(k, kl, prec, &pr);
  if (!err.empty()) return fail(AOA_ERR_UNSUPPORTED, err);
  if (!out) return fail(AOA_ERR_ARG, "null output");
  if (n < (int)pr.size()) return fail(AOA_ERR_ARG, "output too small");
  std::memcpy(out, pr.data(), pr.size() * sizeof(uint64_t));
  return AOA_OK;
}

extern "C" int aoa_moddown(const aoa_params* h, void* stream, uint64_t* d_q_out, const uint64_t* d_qb_in,
                           int level) {
  if (!h || !d_q_out || !d_qb_in) return fail(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  ModDownArgs a;
  std::memset(&a, 0, sizeof(a));
  a.inQ[0] = d_qb_in;
  a.inB[0] = d_qb_in + (size_t)h->host.k * h->host.N;
  a.out[0] = d_q_out;
  a.npolys = 1;
  a.level = level;
  a.rescale = 0;
  launch_moddown(h->dev, a, (hipStream_t)stream);
  return check_launch("aoa_moddown");
}

extern "C" int aoa_automorphism(const aoa_params* h, void* stream, uint64_t* out, const uint64_t* in, int nres,
                                int rot_slots) {
  if (!h || !out || !in) return fail(AOA_ERR_ARG, "null argument");
  if (nres <= 0 || nres > h->host.k + h->host.K) return fail(AOA_ERR_ARG, "limb count out of range");
  if (rot_slots < 0) return fail(AOA_ERR_ARG, "negative rotation");
  launch_automorphism(h->dev, out, in, nres, galois_of(rot_slots, h->host.N), (hipStream_t)stream);
  return check_launch("aoa_automorphism");
}

extern "C" size_t aoa_workspace_bytes(const aoa_params* h, int level) {
  if (!h || level < 0 || level >= h->host.k) return 0;
  const size_t L = (size_t)level + 1, K = (size_t)h->host.K;
  return (6 * L + 3 * K) * (size_t)h->host.N * sizeof(uint64_t);
}

// Key switch of d (coefficient domain, L limbs) into (out0 [+ add0], out1).
// Workspace: X (L+K limbs), P0, P1 (L+K limbs each).
static int keyswitch_core(const aoa_params* h, hipStream_t st, uint64_t* out0, uint64_t* out1,
                          const uint64_t* add0, const uint64_t* d, const uint64_t* kb, const uint64_t* ka,
                          int level, uint64_t* ws) {
  const DevParams& P = h->dev;
  const size_t N = (size_t)P.N;
  const int L = level + 1, K = P.K;
  uint64_t* X = ws;
  uint64_t* P0 = X + (size_t)(L + K) * N;
  uint64_t* P1 = P0 + (size_t)(L + K) * N;
  // ModUp: special limbs of X from d
  launch_modup(P, d, X + (size_t)L * N, level, st);
  // X_Q = NTT(d) (out of place), X_B = NTT(X_B)
  LimbMap mq;
  std::memset(&mq, 0, sizeof(mq));
  mq.base[0] = X;
  mq.src[0] = d;
  mq.nlimbs = L;
  mq.qlimbs = L;
  mq.npolys = 1;
  if (ntt_forward(P, mq, 0, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split");
  LimbMap mb;
  std::memset(&mb, 0, sizeof(mb));
  mb.base[0] = X + (size_t)L * N;
  mb.nlimbs = K;
  mb.qlimbs = 0;
  mb.npolys = 1;
  if (ntt_forward(P, mb, 0, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split");
  // P_y = INTT(X * key_y)
  LimbMap mk;
  std::memset(&mk, 0, sizeof(mk));
  mk.base[0] = P0;
  mk.base[1] = P1;
  mk.src[0] = X;
  mk.nlimbs = L + K;
  mk.qlimbs = L;
  mk.npolys = 2;
  if (ntt_inverse(P, mk, 0, kb, ka, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split");
  ModDownArgs a;
  std::memset(&a, 0, sizeof(a));
  a.inQ[0] = P0;
  a.inQ[1] = P1;
  a.inB[0] = P0 + (size_t)L * N;
  a.inB[1] = P1 + (size_t)L * N;
  a.add[0] = add0;
  a.out[0] = out0;
  a.out[1] = out1;
  a.npolys = 2;
  a.level = level;
  a.rescale = 0;
  launch_moddown(P, a, st);
  return AOA_OK;
}

extern "C" int aoa_keyswitch(const aoa_params* h, void* stream, uint64_t* out0, uint64_t* out1,
                             const uint64_t* add0, const uint64_t* d_in, const uint64_t* key_b,
                             const uint64_t* key_a, int level, void* ws, size_t ws_bytes) {
  if (!h || !out0 || !out1 || !d_in || !key_b || !key_a || !ws) return fail(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  if (ws_bytes < aoa_workspace_bytes(h, level)) return fail(AOA_ERR_WORKSPACE, "workspace too small");
  int rc = keyswitch_core(h, (hipStream_t)stream, out0, out1, add0, d_in, key_b, key_a, level, (uint64_t*)ws);
  if (rc) return rc;
  return check_launch("aoa_keyswitch");
}

int aoa::hommul_core(const aoa_params* h, void* stream, uint64_t* c3_0, uint64_t* c3_1, const uint64_t* a0,
                     const uint64_t* a1, const uint64_t* b0, const uint64_t* b1, const uint64_t* evk_b,
                     const uint64_t* evk_a, int level, void* ws, size_t ws_bytes, int rescale) {
  if (!h || !c3_0 || !c3_1 || !a0 || !a1 || !b0 || !b1 || !evk_b || !evk_a || !ws)
    return fail(AOA_ERR_ARG, "null argument");
  if (!aligned16(c3_0) || !aligned16(c3_1) || !aligned16(a0) || !aligned16(a1) || !aligned16(b0) ||
      !aligned16(b1) || !aligned16(ws))
    return fail(AOA_ERR_ARG, "device pointers must be 16-byte aligned");
  if (level == 0) return fail(AOA_ERR_LEVEL, "The ciphertext level is 0. No more rescaling is possible.");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  if (ws_bytes < aoa_workspace_bytes(h, level)) return fail(AOA_ERR_WORKSPACE, "workspace too small");
  const DevParams& P = h->dev;
  hipStream_t st = (hipStream_t)stream;
  const size_t N = (size_t)P.N;
  const int L = level + 1, K = P.K;
  uint64_t* X = (uint64_t*)ws;            // L+K: d2 (NTT) | special limbs
  uint64_t* D0 = X + (size_t)(L + K) * N;  // L
  uint64_t* D1 = D0 + (size_t)L * N;       // L
  uint64_t* D2 = D1 + (size_t)L * N;       // L: d2 in the coefficient domain
  uint64_t* P0 = D2 + (size_t)L * N;       // L+K
  uint64_t* P1 = P0 + (size_t)(L + K) * N; // L+K
  // With two-pass transforms d0 and d1 stay in the NTT domain and join the
  // key products as (B mod q) d0, (B mod q) d1: ModDown is linear in its Q
  // limbs and its special-limb sum does not see them, so
  // ModDown(X + B d) = ModDown(X) + d residue for residue, and only d2 is
  // transformed back.  The single-workgroup transform (N <= 4096) keeps the
  // additions in ModDown.
  const bool fold = (P.ntt_n1 ? P.ntt_n1 : ntt_default_n1(P.N)) != P.N;
  launch_tensor(P, D0, D1, X, a0, a1, b0, b1, L, st);
  LimbMap mi;
  std::memset(&mi, 0, sizeof(mi));
  if (fold) {
    mi.base[0] = D2;
    mi.src[0] = X;
    mi.npolys = 1;
  } else {
    mi.base[0] = D0;
    mi.base[1] = D1;
    mi.base[2] = D2;
    mi.src[0] = D0;
    mi.src[1] = D1;
    mi.src[2] = X;
    mi.npolys = 3;
  }
  mi.nlimbs = L;
  mi.qlimbs = L;
  if (ntt_inverse(P, mi, 0, nullptr, nullptr, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split");
  launch_modup(P, D2, X + (size_t)L * N, level, st);
  LimbMap mb;
  std::memset(&mb, 0, sizeof(mb));
  mb.base[0] = X + (size_t)L * N;
  mb.nlimbs = K;
  mb.qlimbs = 0;
  mb.npolys = 1;
  if (ntt_forward(P, mb, 0, st)) return fail(AOA_ERR_UNSUPPORTED, "no NTT split");
  LimbMap mk;
  std::memset(&mk, 0, sizeof(mk));
  mk.base[0] = P0;
  mk.base[1] = P1;
  mk.src[0] = X;
  mk.nlimbs = L + K;
  mk.qlimbs = L;
  mk.npolys = 2;
  if (ntt_inverse(P, mk, 0, evk_b, evk_a, st, fold ? D0 : nullptr, fold ? D1 : nullptr))
    return fail(AOA_ERR_UNSUPPORTED, "no NTT split");
  ModDownArgs a;
  std::memset(&a, 0, sizeof(a));
  a.inQ[0] = P0;
  a.inQ[1] = P1;
  a.inB[0] = P0 + (size_t)L * N;
  a.inB[1] = P1 + (size_t)L * N;
  a.add[0] = fold ? nullptr : D0;
  a.add[1] = fold ? nullptr : D1;
  a.out[0] = c3_0;
  a.out[1] = c3_1;
  a.npolys = 2;
  a.level = level;
  a.rescale = rescale;
  launch_moddown(P, a, st);
  return check_launch(rescale ? "aoa_hommul_relin_rescale" : "aoa_hommul_relin");
}

extern "C" int aoa_hommul_relin_rescale(const aoa_params* h, void* stream, uint64_t* c3_0, uint64_t* c3_1,
                                        const uint64_t* a0, const uint64_t* a1, const uint64_t* b0,
                                        const uint64_t* b1, const uint64_t* evk_b, const uint64_t* evk_a,
                                        int level, void* ws, size_t ws_bytes) {
  return aoa::hommul_core(h, stream, c3_0, c3_1, a0, a1, b0, b1, evk_b, evk_a, level, ws, ws_bytes, 1);
}

extern "C" int aoa_hommul_relin(const aoa_params* h, void* stream, uint64_t* c3_0, uint64_t* c3_1,
                                const uint64_t* a0, const uint64_t* a1, const uint64_t* b0, const uint64_t* b1,
                                const uint64_t* evk_b, const uint64_t* evk_a, int level, void* ws,
                                size_t ws_bytes) {
  return aoa::hommul_core(h, stream, c3_0, c3_1, a0, a1, b0, b1, evk_b, evk_a, level, ws, ws_bytes, 0);
}

// rho (aoa-ntt/src/cuda/manager.cu:149-176): Q limbs times B mod q_i, special
// limbs zero.  Any domain (the factor is a scalar per limb).
extern "C" int aoa_rho(const aoa_params* h, void* stream, uint64_t* out, const uint64_t* in) {
  if (!h || !out || !in) return fail(AOA_ERR_ARG, "null argument");
  launch_rho(h->dev, out, in, (hipStream_t)stream);
  return check_launch("aoa_rho");
}

extern "C" int aoa_rotate_left(const aoa_params* h, void* stream, uint64_t* out0, uint64_t* out1,
                               const uint64_t* c0, const uint64_t* c1, const uint64_t* rtk_b,
                               const uint64_t* rtk_a, int rot_slots, int level, void* ws, size_t ws_bytes) {
  if (!h || !out0 || !out1 || !c0 || !c1 || !rtk_b || !rtk_a || !ws) return fail(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return fail(AOA_ERR_ARG, "level out of range");
  if (rot_slots < 0) return fail(AOA_ERR_ARG, "negative rotation");
  if (ws_bytes < aoa_workspace_bytes(h, level)) return fail(AOA_ERR_WORKSPACE, "workspace too small");
  const DevParams& P = h->dev;
  hipStream_t st = (hipStream_t)stream;
  const size_t N = (size_t)P.N;
  const int L = level + 1, K = P.K;
  uint64_t* R0 = (uint64_t*)ws;
  uint64_t* R1 = R0 + (size_t)L * N;
  uint64_t* kws = R1 + (size_t)L * N;  // 3(L+K) limbs
  (void)K;
  const uint64_t g = galois_of(rot_slots, P.N);
  launch_automorphism(P, R0, c0, L, g, st);
  launch_automorphism(P, R1, c1, L, g, st);
  int rc = keyswitch_core(h, st, out0, out1, R0, R1, rtk_b, rtk_a, level, kws);
  if (rc) return rc;
  return check_launch("aoa_rotate_left");
}

// ckks_rotate_right (aoa-ntt/src/ckks/ckks.cpp:953-980): rotate_slots_right is
// rotate_slots_left by N/2 - rotSlots (aoa-ntt/src/ckks/encoder.cu:356-363);
// the key switch is the same, with the right rotation key.
extern "C" int aoa_rotate_right(const aoa_params* h, void* stream, uint64_t* out0, uint64_t* out1,
                                const uint64_t* c0, const uint64_t* c1, const uint64_t* rtk_b,
                                const uint64_t* rtk_a, int rot_slots, int level, void* ws, size_t ws_bytes) {
  if (!h) return fail(AOA_ERR_ARG, "null argument");
  if (rot_slots < 0 || rot_slots > h->host.N / 2) return fail(AOA_ERR_ARG, "rotation out of range [0, N/2]");
  return aoa_rotate_left(h, stream, out0, out1, c0, c1, rtk_b, rtk_a, h->host.N / 2 - rot_slots, level, ws,
                         ws_bytes);
}
