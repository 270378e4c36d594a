// Batched (stage-fused) evaluation: many ciphertexts, mixed levels and ring
// dimensions, one launch per pipeline stage per parameter set, the whole
// sequence captured once as a hipGraph and replayed.
//
// A batch is a list of layers; a layer holds independent operations
// (HomMul+relin+rescale, rotate[-and-accumulate], forward/inverse NTT) grouped
// by parameter set.  Layers run in order (each may consume the previous one's
// outputs, as the rotate-and-sum steps of ckks_sumslots do,
// aoa-ntt/src/ckks/ckks.cpp:479-495).  Every stage of a group is one
// descriptor-driven kernel launch covering all its ciphertexts, so a launch
// carries thousands of workgroups even when each ciphertext is small, and
// evaluation keys are shared between them.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace aoa;

namespace aoa {

struct BatchGroup {
  const aoa_params* params = nullptr;
  const uint64_t* evk_b = nullptr;
  const uint64_t* evk_a = nullptr;
  std::vector<aoa_mul_item> muls;
  std::vector<aoa_rot_item> rots;
  std::vector<aoa_ntt_item> ntts;
  // descriptor tables (device) and counts, set by batch_fill
  NttDesc* d_pre_fwd = nullptr;
  int n_pre_fwd = 0;
  NttDesc* d_pre_inv = nullptr;
  int n_pre_inv = 0;
  TensorDesc* d_tensor = nullptr;
  int n_tensor = 0;
  AutoDesc* d_auto = nullptr;
  int n_auto = 0;
  NttDesc* d_intt3 = nullptr;  // mul: INTT of d0, d1, d2
  int n_intt3 = 0;
  NttDesc* d_trow = nullptr;  // mul (N >= 8192): inverse NTT of d2 = a1 b1, the product in its first pass
  int n_trow = 0, ns_trow = -1;
  ModUpDesc* d_modup = nullptr;
  int n_modup = 0;
  int max_level = 0;
  NttDesc* d_fwd = nullptr;  // NTT of ModUp output (and of the rotated c1)
  int n_fwd = 0;
  KeyDesc2* d_key2 = nullptr;  // key product + first INTT pass, both key halves
  int n_key2 = 0;
  NttDesc* d_keycol = nullptr;  // second INTT pass of the key products
  int n_keycol = 0;
  ModDownDesc* d_moddown = nullptr;
  int n_moddown = 0;
  // leading descriptors whose prime is 2^61 or above (-1: not a prefix)
  int ns_key2 = -1;
};

struct BatchLayer {
  std::vector<BatchGroup> groups;
};

// Workspace limbs of one HomMul item: X (L + K), d0, d1, d2 (L each), P0, P1
// (L + K each); from N = 8192 up (fused) the tensor product is never stored
// and X's Q limbs are not needed: X's special limbs (K), d2 (L), P0, P1.
size_t mul_ws_limbs(int L, int K, bool fused) { return (fused ? 3 : 6) * (size_t)L + 3 * (size_t)K; }
size_t rot_ws_limbs(int L, int K) { return 5 * (size_t)L + 3 * (size_t)K; }

}  // namespace aoa

struct aoa_batch {
  std::vector<BatchLayer> layers;
  aoa_batch_options opt{0, 1, 0, 0};
  bool built = false;
  void* dmem = nullptr;  // descriptors
  void* ws = nullptr;    // workspace
  size_t ws_bytes = 0;
  size_t desc_bytes = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BatchLayer* layers_cur = nullptr;  // layer being enqueued
};

namespace aoa {

static BatchGroup& group_for(aoa_batch* b, const aoa_params* p, const uint64_t* kb, const uint64_t* ka) {
  if (b->layers.empty()) b->layers.emplace_back();
  BatchLayer& L = b->layers.back();
  for (auto& g : L.groups)
    if (g.params == p && g.evk_b == kb && g.evk_a == ka) return g;
  L.groups.emplace_back();
  BatchGroup& g = L.groups.back();
  g.params = p;
  g.evk_b = kb;
  g.evk_a = ka;
  return g;
}

// Key-switching product descriptors of one item: X (Lq Q limbs at XQ, K
// special limbs at XB, NTT domain) times both key halves into P0 / P1, one
// KeyDesc2 per limb for the first inverse pass and one NttDesc per output
// limb for the second.  A0 / A1 (Lq limbs, NTT domain, or null) join the Q
// limbs of P0 / P1 times B mod q, so that ModDown adds them (two-pass
// transforms only).  With c1 / c2 set (HomMul's operands, two-pass
// transforms) the Q limbs take X = a1 b1 and the adds d0, d1 from the
// operands in the pass instead (KeyDesc2::ab), and XQ is not read (null).
static void add_key_descs(std::vector<KeyDesc2>& key2, std::vector<NttDesc>& keycol, const uint64_t* XQ,
                          const uint64_t* XB, uint64_t* P0, uint64_t* P1, const uint64_t* kb, const uint64_t* ka,
                          const uint64_t* A0, const uint64_t* A1, int Lq, int K, int k, size_t N,
                          const uint64_t* const* c1 = nullptr, const uint64_t* const* c2 = nullptr) {
  for (int m = 0; m < Lq + K; ++m) {
    const int pr = m < Lq ? m : k + (m - Lq);
    const size_t o = (size_t)m * N;
    KeyDesc2 e;
    e.src = m < Lq ? (XQ ? XQ + o : nullptr) : XB + (size_t)(m - Lq) * N;
    e.dst[0] = P0 + o;
    e.dst[1] = P1 + o;
    e.key[0] = kb + (size_t)pr * N;
    e.key[1] = ka + (size_t)pr * N;
    e.add[0] = A0 && m < Lq ? A0 + o : nullptr;
    e.add[1] = A1 && m < Lq ? A1 + o : nullptr;
    const bool t = c1 && c2 && m < Lq;
    e.ab[0] = t ? c1[0] + o : nullptr;
    e.ab[1] = t ? c1[1] + o : nullptr;
    e.ab[2] = t ? c2[0] + o : nullptr;
    e.ab[3] = t ? c2[1] + o : nullptr;
    e.prime = pr;
    e.pad = 0;
    key2.push_back(e);
  }
  for (int y = 0; y < 2; ++y)
    for (int m = 0; m < Lq + K; ++m) {
      uint64_t* d = (y ? P1 : P0) + (size_t)m * N;
      keycol.push_back({d, d, nullptr, m < Lq ? m : k + (m - Lq), 0});
    }
}

// Workspace bytes (max over layers of the layer's sum) and descriptor bytes.
void batch_sizes(const aoa_batch* b, size_t* ws_bytes, size_t* desc_bytes) {
  size_t ws = 0, desc = 0;
  for (auto& L : b->layers) {
    size_t s = 0;
    for (auto& g : L.groups) {
      const size_t N = g.params->host.N, K = g.params->host.K;
      for (auto& it : g.muls) {
        const size_t Lq = it.level + 1;
        s += mul_ws_limbs((int)Lq, (int)K, N >= 8192) * N * 8;
        desc += Lq * (sizeof(TensorDesc) + 3 * sizeof(NttDesc)) + sizeof(ModUpDesc) +
                K * sizeof(NttDesc) +
                (Lq + K) * sizeof(KeyDesc2) + 2 * (Lq + K) * sizeof(NttDesc) + 2 * sizeof(ModDownDesc);
      }
      for (auto& it : g.rots) {
        const size_t Lq = it.level + 1;
        s += rot_ws_limbs((int)Lq, (int)K) * N * 8;
        desc += Lq * (2 * sizeof(AutoDesc) + sizeof(NttDesc)) + sizeof(ModUpDesc) + K * sizeof(NttDesc) +
                (Lq + K) * sizeof(KeyDesc2) + 2 * (Lq + K) * sizeof(NttDesc) + 2 * sizeof(ModDownDesc);
      }
      for (auto& it : g.ntts) desc += 2 * (size_t)(it.level + 1) * sizeof(NttDesc);
      desc += 16 * 256;  // per-table alignment slack
    }
    ws = std::max(ws, s);
  }
  *ws_bytes = ws;
  *desc_bytes = desc;
}

// Fills every descriptor table into `host` (desc_cap bytes) using workspace
// `ws` and device descriptor base `dev`; device table pointers are set to
// dev + offset.  Returns the bytes used, or 0 if desc_cap is too small.
size_t batch_fill(aoa_batch* b, uint64_t* ws, char* host, char* dev, size_t desc_cap) {
  size_t used = 0;
  bool overflow = false;
  auto put = [&](const void* p, size_t bytes) -> char* {
    size_t off = (used + 255) & ~(size_t)255;
    if (off + bytes > desc_cap) {
      overflow = true;
      return nullptr;
    }
    std::memcpy(host + off, p, bytes);
    used = off + bytes;
    return dev + off;
  };
  for (auto& L : b->layers) {
    uint64_t* w = ws;  // layers reuse the workspace
    for (auto& g : L.groups) {
      const DevParams& P = g.params->dev;
      const size_t N = (size_t)P.N;
      const int k = P.k, K = P.K;
      std::vector<TensorDesc> tens;
      std::vector<NttDesc> prefwd, preinv, intt3, fwd, keycol, trow;
      std::vector<KeyDesc2> key2;
      std::vector<ModUpDesc> mu;
      std::vector<ModDownDesc> md;
      std::vector<AutoDesc> au;
      g.max_level = 0;
      for (auto& it : g.ntts)
        for (int y = 0; y < 2; ++y)
          for (int i = 0; i <= it.level; ++i) {
            uint64_t* l = it.poly[y] + (size_t)i * N;
            (it.dir == AOA_FORWARD ? prefwd : preinv).push_back({l, l, nullptr, i, 0});
          }
      for (auto& it : g.muls) {
        const int lv = it.level, Lq = lv + 1;
        g.max_level = std::max(g.max_level, lv);
        const bool fused = P.N >= 8192;
        // unfused: X (Lq + K), D0, D1, D2, P0, P1; fused: X's special limbs, D2, P0, P1
        uint64_t* X = fused ? nullptr : w;
        uint64_t* XB = fused ? w : X + (size_t)Lq * N;
        uint64_t* D0 = fused ? nullptr : XB + (size_t)K * N;
        uint64_t* D1 = fused ? nullptr : D0 + (size_t)Lq * N;
        uint64_t* D2 = fused ? XB + (size_t)K * N : D1 + (size_t)Lq * N;
        uint64_t* P0 = D2 + (size_t)Lq * N;
        uint64_t* P1 = P0 + (size_t)(Lq + K) * N;
        w += mul_ws_limbs(Lq, K, fused) * N;
        // From N = 8192 up the transforms are two-pass and the tensor product
        // is never stored: d2 = a1 b1 is formed in the first inverse pass of
        // d2 (for ModUp), and again with d0 and d1 in the key products' first
        // pass, where d0, d1 join the Q limbs as B d0, B d1: ModDown is linear
        // in its Q limbs and its B-limb sum does not see them, so
        // ModDown(X + B d) = ModDown(X) + d, residue for residue, and d0, d1
        // need no transform.  Below, a tensor kernel, the single-workgroup
        // transforms, and the additions in ModDown.
        for (int i = 0; i < Lq && fused; ++i) {
          const size_t o = (size_t)i * N;
          trow.push_back({it.c1[1] + o, D2 + o, it.c2[1] + o, i, 0});
        }
        for (int i = 0; i < Lq && !fused; ++i) {
          const size_t o = (size_t)i * N;
          TensorDesc t;
          t.a0 = it.c1[0] + o;
          t.a1 = it.c1[1] + o;
          t.b0 = it.c2[0] + o;
          t.b1 = it.c2[1] + o;
          t.d0 = D0 + o;
          t.d1 = D1 + o;
          t.d2 = X + o;
          t.prime = i;
          t.pad = 0;
          tens.push_back(t);
          intt3.push_back({D0 + o, D0 + o, nullptr, i, 0});
          intt3.push_back({D1 + o, D1 + o, nullptr, i, 0});
          intt3.push_back({X + o, D2 + o, nullptr, i, 0});
        }
        mu.push_back({D2, XB, lv, 0});
        for (int j = 0; j < K; ++j) fwd.push_back({XB + (size_t)j * N, XB + (size_t)j * N, nullptr, k + j, 0});
        add_key_descs(key2, keycol, X, XB, P0, P1, g.evk_b, g.evk_a, nullptr, nullptr, Lq, K, k, N,
                      fused ? it.c1 : nullptr, fused ? it.c2 : nullptr);
        for (int y = 0; y < 2; ++y)
          md.push_back({y ? P1 : P0, (y ? P1 : P0) + (size_t)Lq * N, fused ? nullptr : (y ? D1 : D0), it.out[y],
                        lv, 1});
      }
      for (auto& it : g.rots) {
        const int lv = it.level, Lq = lv + 1;
        g.max_level = std::max(g.max_level, lv);
        uint64_t* R0 = w;
        uint64_t* R1 = R0 + (size_t)Lq * N;
        uint64_t* X = R1 + (size_t)Lq * N;
        uint64_t* P0 = X + (size_t)(Lq + K) * N;
        uint64_t* P1 = P0 + (size_t)(Lq + K) * N;
        w += rot_ws_limbs(Lq, K) * N;
        const uint64_t ginv = galois_inverse(galois_of(it.rot_slots, P.N), P.N);
        for (int i = 0; i < Lq; ++i) {
          const size_t o = (size_t)i * N;
          au.push_back({it.in[0] + o, R0 + o, it.accumulate ? it.in[0] + o : nullptr, ginv, i, 0});
          au.push_back({it.in[1] + o, R1 + o, nullptr, ginv, i, 0});
          fwd.push_back({R1 + o, X + o, nullptr, i, 0});
        }
        mu.push_back({R1, X + (size_t)Lq * N, lv, 0});
        for (int j = 0; j < K; ++j)
          fwd.push_back({X + (size_t)(Lq + j) * N, X + (size_t)(Lq + j) * N, nullptr, k + j, 0});
        add_key_descs(key2, keycol, X, X + (size_t)Lq * N, P0, P1, it.rtk_b, it.rtk_a, nullptr, nullptr, Lq, K, k, N);
        for (int y = 0; y < 2; ++y) {
          const uint64_t* add = y == 0 ? R0 : (it.accumulate ? it.in[1] : nullptr);
          md.push_back({y ? P1 : P0, (y ? P1 : P0) + (size_t)Lq * N, add, it.out[y], lv, 0});
        }
      }
      // Prime-major order: the limbs of one prime from every item run next
      // to each other, so their workgroups (spread over the XCDs by block
      // index) find that prime's twiddles, and in the key stage its key
      // limbs, in the L2 cache.  Entries of a stage are independent.
      // Item order (prime_major = 0) keeps the order of insertion: the limbs of
      // one item after the other.
      const bool pm = b->opt.prime_major != 0;
      auto by_prime = [pm](auto& v) {
        if (pm) std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.prime < b.prime; });
      };
      by_prime(prefwd);
      by_prime(preinv);
      by_prime(intt3);
      by_prime(fwd);
      by_prime(key2);
      by_prime(keycol);
      by_prime(trow);
      // The butterflies of primes of 2^61 and above (the 63-bit q0) are not
      // lazy; they sort first, and the key stage's first pass over the other
      // limbs is launched in its lazy-only form (98 instead of 160 registers,
      // four waves per SIMD instead of three).  The other stages gain nothing
      // from the split (their occupancy is bound by LDS) and keep one launch.
      auto strict_prefix = [&](const auto& v) -> int {
        const std::vector<uint64_t>& pr = g.params->host.primes;
        int n = 0;
        while (n < (int)v.size() && pr[v[n].prime] >= (1ull << 61)) ++n;
        for (size_t t = n; t < v.size(); ++t)
          if (pr[v[t].prime] >= (1ull << 61)) return -1;
        return n;
      };
      // Small stages (a serving step of a few requests) keep one launch by
      // default: the split pays only where a launch fills the GPU.  The split
      // needs the q0 limbs first, i.e. prime-major order.
      constexpr size_t kSplitMin = 256;
      auto split_at = [&](const auto& v) -> int {
        if (b->opt.strict_split < 0 || !pm) return -1;
        if (b->opt.strict_split == 0 && v.size() < kSplitMin) return -1;
        return strict_prefix(v);
      };
      g.ns_key2 = split_at(key2);
      g.ns_trow = split_at(trow);
      g.n_trow = (int)trow.size();
      g.d_trow = (NttDesc*)(trow.empty() ? nullptr : put(trow.data(), trow.size() * sizeof(NttDesc)));
      g.n_pre_fwd = (int)prefwd.size();
      g.n_pre_inv = (int)preinv.size();
      g.n_tensor = (int)tens.size();
      g.n_intt3 = (int)intt3.size();
      g.n_auto = (int)au.size();
      g.n_fwd = (int)fwd.size();
      g.n_modup = (int)mu.size();
      g.n_key2 = (int)key2.size();
      g.n_keycol = (int)keycol.size();
      g.n_moddown = (int)md.size();
      g.d_pre_fwd = (NttDesc*)(prefwd.empty() ? nullptr : put(prefwd.data(), prefwd.size() * sizeof(NttDesc)));
      g.d_pre_inv = (NttDesc*)(preinv.empty() ? nullptr : put(preinv.data(), preinv.size() * sizeof(NttDesc)));
      g.d_tensor = (TensorDesc*)(tens.empty() ? nullptr : put(tens.data(), tens.size() * sizeof(TensorDesc)));
      g.d_intt3 = (NttDesc*)(intt3.empty() ? nullptr : put(intt3.data(), intt3.size() * sizeof(NttDesc)));
      g.d_auto = (AutoDesc*)(au.empty() ? nullptr : put(au.data(), au.size() * sizeof(AutoDesc)));
      g.d_fwd = (NttDesc*)(fwd.empty() ? nullptr : put(fwd.data(), fwd.size() * sizeof(NttDesc)));
      g.d_modup = (ModUpDesc*)(mu.empty() ? nullptr : put(mu.data(), mu.size() * sizeof(ModUpDesc)));
      g.d_key2 = (KeyDesc2*)(key2.empty() ? nullptr : put(key2.data(), key2.size() * sizeof(KeyDesc2)));
      g.d_keycol = (NttDesc*)(keycol.empty() ? nullptr : put(keycol.data(), keycol.size() * sizeof(NttDesc)));
      g.d_moddown = (ModDownDesc*)(md.empty() ? nullptr : put(md.data(), md.size() * sizeof(ModDownDesc)));
    }
  }
  return overflow ? 0 : used;
}

// Records a HIP event pair around each stage launch of batch_enqueue (on the
// stream the stage is launched on); used only by aoa_batch_stage_times.
struct StageTimer {
  std::vector<int> stage;
  std::vector<hipEvent_t> ev;  // begin, end pairs
  bool ok = true;
  void begin(int s, hipStream_t st) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) { ok = false; return; }
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); ok = false; return; }
    stage.push_back(s);
    ev.push_back(e0);
    ev.push_back(e1);
    (void)hipEventRecord(e0, st);
  }
  void end(hipStream_t st) {
    if (ok) (void)hipEventRecord(ev.back(), st);
  }
  ~StageTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// Issues every stage launch of the batch on stream st (stage order
// AOA_STAGE_*); with a timer, each (layer, stage) launch run is bracketed by
// an event pair.
// The parameters a group's launches use: its handle's, with the batch's NTT
// split (when the group's ring admits it) and XCD mapping.
static DevParams launch_params(const aoa_batch* b, const BatchGroup& g) {
  DevParams P = g.params->dev;
  if (b->opt.ntt_split && ntt_valid_split(P.N, b->opt.ntt_split)) P.ntt_n1 = b->opt.ntt_split;
  P.xcd_chunks = b->opt.xcd_chunks ? 1 : 0;
  return P;
}

int batch_enqueue(aoa_batch* b, hipStream_t st, StageTimer* tm) {
  int rc = AOA_OK;
  auto run = [&](int s, auto&& has, auto&& go) {
    bool any = false;
    for (auto& g : b->layers_cur->groups) any = any || has(g);
    if (!any || rc) return;
    if (tm) tm->begin(s, st);
    for (auto& g : b->layers_cur->groups)
      if (has(g) && go(g)) rc = set_error(AOA_ERR_UNSUPPORTED, "no NTT split");
    if (tm) tm->end(st);
  };
  for (auto& L : b->layers) {
    b->layers_cur = &L;
    run(AOA_STAGE_PRE_FWD, [](BatchGroup& g) { return g.n_pre_fwd > 0; },
        [&](BatchGroup& g) { return ntt_forward_desc(launch_params(b, g), g.d_pre_fwd, g.n_pre_fwd, 0, st); });
    run(AOA_STAGE_PRE_INV, [](BatchGroup& g) { return g.n_pre_inv > 0; },
        [&](BatchGroup& g) { return ntt_inverse_desc(launch_params(b, g), g.d_pre_inv, g.n_pre_inv, 0, st); });
    run(AOA_STAGE_TENSOR, [](BatchGroup& g) { return g.n_tensor > 0; }, [&](BatchGroup& g) {
      launch_tensor_desc(launch_params(b, g), g.d_tensor, g.n_tensor, st);
      return 0;
    });
    run(AOA_STAGE_AUTOMORPHISM, [](BatchGroup& g) { return g.n_auto > 0; }, [&](BatchGroup& g) {
      launch_automorphism_desc(launch_params(b, g), g.d_auto, g.n_auto, st);
      return 0;
    });
    run(AOA_STAGE_INTT, [](BatchGroup& g) { return g.n_intt3 > 0 || g.n_trow > 0; }, [&](BatchGroup& g) {
      // d2 goes to ModUp only, whose Shoup products take [0, 2p)
      if (g.n_trow > 0 && ntt_inverse_desc(launch_params(b, g), g.d_trow, g.n_trow, 0, st, g.ns_trow, false))
        return -1;
      return g.n_intt3 > 0 ? ntt_inverse_desc(launch_params(b, g), g.d_intt3, g.n_intt3, 0, st) : 0;
    });
    run(AOA_STAGE_MODUP, [](BatchGroup& g) { return g.n_modup > 0; }, [&](BatchGroup& g) {
      launch_modup_desc(launch_params(b, g), g.d_modup, g.n_modup, g.max_level, st);
      return 0;
    });
    run(AOA_STAGE_NTT, [](BatchGroup& g) { return g.n_fwd > 0; },
        [&](BatchGroup& g) { return ntt_forward_desc(launch_params(b, g), g.d_fwd, g.n_fwd, 0, st); });
    run(AOA_STAGE_KEY_INTT, [](BatchGroup& g) { return g.n_key2 > 0; }, [&](BatchGroup& g) {
      // the key products go to ModDown only, which takes [0, 2p) except for q0
      return ntt_inverse_key2_desc(launch_params(b, g), g.d_key2, g.n_key2, g.d_keycol, g.n_keycol, 0, st, g.ns_key2,
                                   -1, false);
    });
    run(AOA_STAGE_MODDOWN, [](BatchGroup& g) { return g.n_moddown > 0; }, [&](BatchGroup& g) {
      // the key stage above (canon = false) hands over special limbs
      // already multiplied by (B/b_j)^-1 in its two-pass form
      const DevParams P = launch_params(b, g);
      launch_moddown_desc(P, g.d_moddown, g.n_moddown, ntt_key2_bscaled(P, 0), st);
      return 0;
    });
  }
  b->layers_cur = nullptr;
  if (rc) return rc;
  return check_launch("aoa_batch");
}

int batch_add_mul(aoa_batch* b, const aoa_params* p, const uint64_t* evk_b, const uint64_t* evk_a,
                  const aoa_mul_item* items, int n) {
  for (int i = 0; i < n; ++i) {
    const aoa_mul_item& it = items[i];
    if (!it.c1[0] || !it.c1[1] || !it.c2[0] || !it.c2[1] || !it.out[0] || !it.out[1])
      return set_error(AOA_ERR_ARG, "null pointer in item");
    if ((((uintptr_t)it.c1[0] | (uintptr_t)it.c1[1] | (uintptr_t)it.c2[0] | (uintptr_t)it.c2[1]) & 15) != 0)
      return set_error(AOA_ERR_ARG, "device pointers must be 16-byte aligned");
    if (it.level == 0) return set_error(AOA_ERR_LEVEL, "The ciphertext level is 0. No more rescaling is possible.");
    if (it.level < 0 || it.level >= p->host.k) return set_error(AOA_ERR_ARG, "level out of range");
  }
  BatchGroup& g = group_for(b, p, evk_b, evk_a);
  g.muls.insert(g.muls.end(), items, items + n);
  return AOA_OK;
}

int batch_add_rotate(aoa_batch* b, const aoa_params* p, const aoa_rot_item* items, int n) {
  for (int i = 0; i < n; ++i) {
    const aoa_rot_item& it = items[i];
    if (!it.in[0] || !it.in[1] || !it.out[0] || !it.out[1] || !it.rtk_b || !it.rtk_a)
      return set_error(AOA_ERR_ARG, "null pointer in item");
    if (it.level < 0 || it.level >= p->host.k || it.rot_slots < 0) return set_error(AOA_ERR_ARG, "bad item");
  }
  BatchGroup& g = group_for(b, p, nullptr, nullptr);
  g.rots.insert(g.rots.end(), items, items + n);
  return AOA_OK;
}

int batch_add_ntt(aoa_batch* b, const aoa_params* p, const aoa_ntt_item* items, int n) {
  for (int i = 0; i < n; ++i) {
    const aoa_ntt_item& it = items[i];
    if (!it.poly[0] || !it.poly[1]) return set_error(AOA_ERR_ARG, "null pointer in item");
    if (it.level < 0 || it.level >= p->host.k) return set_error(AOA_ERR_ARG, "level out of range");
    if (it.dir != AOA_FORWARD && it.dir != AOA_INVERSE) return set_error(AOA_ERR_ARG, "bad direction");
  }
  BatchGroup& g = group_for(b, p, nullptr, nullptr);
  g.ntts.insert(g.ntts.end(), items, items + n);
  return AOA_OK;
}

void batch_clear(aoa_batch* b) {
  b->layers.clear();
  b->layers.emplace_back();
}

bool batch_empty(const aoa_batch* b) {
  for (auto& L : b->layers)
    for (auto& g : L.groups)
      if (!g.muls.empty() || !g.rots.empty() || !g.ntts.empty()) return false;
  return true;
}

void batch_barrier(aoa_batch* b) {
  if (!b->layers.back().groups.empty()) b->layers.emplace_back();
}

}  // namespace aoa

// ------------------------------------------------------------------ C-ABI --

extern "C" int aoa_batch_create(aoa_batch** out) {
  if (!out) return set_error(AOA_ERR_ARG, "null output handle");
  *out = new aoa_batch();
  (*out)->layers.emplace_back();
  return AOA_OK;
}

extern "C" void aoa_batch_destroy(aoa_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->dmem) (void)hipFree(b->dmem);
  if (b->ws) (void)hipFree(b->ws);
  delete b;
}

extern "C" int aoa_batch_add_mul(aoa_batch* b, const aoa_params* p, const uint64_t* evk_b, const uint64_t* evk_a,
                                 const aoa_mul_item* items, int n) {
  if (!b || !p || !evk_b || !evk_a || (!items && n > 0)) return set_error(AOA_ERR_ARG, "null argument");
  if (b->built) return set_error(AOA_ERR_ARG, "batch already built");
  return batch_add_mul(b, p, evk_b, evk_a, items, n);
}

extern "C" int aoa_batch_add_rotate(aoa_batch* b, const aoa_params* p, const aoa_rot_item* items, int n) {
  if (!b || !p || (!items && n > 0)) return set_error(AOA_ERR_ARG, "null argument");
  if (b->built) return set_error(AOA_ERR_ARG, "batch already built");
  return batch_add_rotate(b, p, items, n);
}

extern "C" int aoa_batch_add_ntt(aoa_batch* b, const aoa_params* p, const aoa_ntt_item* items, int n) {
  if (!b || !p || (!items && n > 0)) return set_error(AOA_ERR_ARG, "null argument");
  if (b->built) return set_error(AOA_ERR_ARG, "batch already built");
  return batch_add_ntt(b, p, items, n);
}

extern "C" int aoa_batch_barrier(aoa_batch* b) {
  if (!b) return set_error(AOA_ERR_ARG, "null batch");
  if (b->built) return set_error(AOA_ERR_ARG, "batch already built");
  batch_barrier(b);
  return AOA_OK;
}

extern "C" int aoa_batch_build(aoa_batch* b) {
  if (!b) return set_error(AOA_ERR_ARG, "null batch");
  if (b->built) return AOA_OK;
  size_t ws = 0, desc = 0;
  batch_sizes(b, &ws, &desc);
  b->ws_bytes = ws;
  if (ws && hipMalloc(&b->ws, ws) != hipSuccess) return set_error(AOA_ERR_HIP, "hipMalloc of batch workspace failed");
  if (desc && hipMalloc(&b->dmem, desc) != hipSuccess) return set_error(AOA_ERR_HIP, "hipMalloc of descriptors failed");
  std::vector<char> host(desc);
  size_t used = batch_fill(b, (uint64_t*)b->ws, host.data(), (char*)b->dmem, desc);
  if (desc && !used) return set_error(AOA_ERR_ARG, "descriptor sizing failed");
  b->desc_bytes = used;
  if (used && hipMemcpy(b->dmem, host.data(), used, hipMemcpyHostToDevice) != hipSuccess)
    return set_error(AOA_ERR_HIP, "hipMemcpy of descriptors failed");
  b->built = true;
  return AOA_OK;
}

extern "C" size_t aoa_batch_workspace_bytes(const aoa_batch* b) { return b ? b->ws_bytes : 0; }

extern "C" void aoa_batch_default_options(aoa_batch_options* o) {
  if (o) *o = aoa_batch_options{0, 1, 0, 0};
}

extern "C" int aoa_batch_get_options(const aoa_batch* b, aoa_batch_options* o) {
  if (!b || !o) return set_error(AOA_ERR_ARG, "null argument");
  *o = b->opt;
  return AOA_OK;
}

static void drop_graph(aoa_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

extern "C" int aoa_batch_set_options(aoa_batch* b, const aoa_batch_options* o) {
  if (!b || !o) return set_error(AOA_ERR_ARG, "null argument");
  if (o->strict_split < -1 || o->strict_split > 1) return set_error(AOA_ERR_ARG, "strict_split must be -1, 0 or 1");
  if (o->ntt_split) {
    bool any = false;
    for (auto& L : b->layers)
      for (auto& g : L.groups) any = any || ntt_valid_split(g.params->host.N, o->ntt_split);
    if (!any) return set_error(AOA_ERR_ARG, "the NTT split fits no ring dimension of the batch");
  }
  b->opt = *o;
  b->opt.prime_major = o->prime_major ? 1 : 0;
  b->opt.xcd_chunks = o->xcd_chunks ? 1 : 0;
  if (!b->built) return AOA_OK;
  // same descriptors in another order: refill the tables in place
  if (hipDeviceSynchronize() != hipSuccess) return set_error(AOA_ERR_HIP, "hipDeviceSynchronize failed");
  drop_graph(b);
  std::vector<char> host(b->desc_bytes);
  size_t used = batch_fill(b, (uint64_t*)b->ws, host.data(), (char*)b->dmem, b->desc_bytes);
  if (b->desc_bytes && used != b->desc_bytes) return set_error(AOA_ERR_ARG, "descriptor refill failed");
  if (used && hipMemcpy(b->dmem, host.data(), used, hipMemcpyHostToDevice) != hipSuccess)
    return set_error(AOA_ERR_HIP, "hipMemcpy of descriptors failed");
  return AOA_OK;
}

extern "C" int aoa_batch_launch(aoa_batch* b, void* stream, int use_graph) {
  if (!b) return set_error(AOA_ERR_ARG, "null batch");
  int rc = aoa_batch_build(b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return batch_enqueue(b, st, nullptr);
  if (!b->exec) {
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipStreamCreate failed");
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipStreamBeginCapture failed");
    int erc = batch_enqueue(b, b->cap, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (erc) {
      if (g) (void)hipGraphDestroy(g);
      return erc;
    }
    if (e != hipSuccess || !g) return set_error(AOA_ERR_HIP, "hipStreamEndCapture failed");
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipGraphInstantiate failed");
  }
  if (hipGraphLaunch(b->exec, st) != hipSuccess) return set_error(AOA_ERR_HIP, "hipGraphLaunch failed");
  return AOA_OK;
}

extern "C" int aoa_batch_stage_times(aoa_batch* b, void* stream, double* ms, int* launches, int nstages) {
  if (!b || !ms || !launches || nstages < AOA_NSTAGES) return set_error(AOA_ERR_ARG, "bad argument");
  int rc = aoa_batch_build(b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  StageTimer tm;
  rc = batch_enqueue(b, st, &tm);
  if (!tm.ok) return set_error(AOA_ERR_HIP, "hipEventCreate failed");
  if (rc) return rc;
  if (hipStreamSynchronize(st) != hipSuccess) return set_error(AOA_ERR_HIP, "hipStreamSynchronize failed");
  for (int s = 0; s < nstages; ++s) {
    ms[s] = 0.0;
    launches[s] = 0;
  }
  for (size_t i = 0; i < tm.stage.size(); ++i) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, tm.ev[2 * i], tm.ev[2 * i + 1]) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipEventElapsedTime failed");
    ms[tm.stage[i]] += t;
    launches[tm.stage[i]] += 1;
  }
  return AOA_OK;
}

// ------------------------------------------------------------ autotuner ----
// The reference fixes its NTT factorisation per ring dimension in a table
// (aoa-ntt/include/newckks/cuda/htrans/ntt.h:13-44).  Here every launch option
// of a batch is timed on the batch itself and the fastest is kept.

// Words of every output of the batch: HomMul outputs after the rescale (level
// limbs), rotation outputs (level + 1 limbs), transformed polynomials.
static void output_ranges(const aoa_batch* b, std::vector<std::pair<const uint64_t*, size_t>>* out) {
  for (auto& L : b->layers)
    for (auto& g : L.groups) {
      const size_t N = g.params->host.N;
      for (auto& it : g.muls)
        for (int y = 0; y < 2; ++y) out->push_back({it.out[y], (size_t)it.level * N});
      for (auto& it : g.rots)
        for (int y = 0; y < 2; ++y) out->push_back({it.out[y], (size_t)(it.level + 1) * N});
      for (auto& it : g.ntts)
        for (int y = 0; y < 2; ++y) out->push_back({it.poly[y], (size_t)(it.level + 1) * N});
    }
}

extern "C" int aoa_batch_autotune(aoa_batch* b, void* stream, int reps, aoa_tune_result* results, int max) {
  if (!b) return -set_error(AOA_ERR_ARG, "null batch");
  if (reps <= 0) reps = 5;
  int rc = aoa_batch_build(b);
  if (rc) return -rc;
  hipStream_t st = (hipStream_t)stream;
  // NTT splits valid for every ring of the batch (0: each handle's own)
  std::vector<int> splits;
  for (int n1 : {64, 128, 256, 2048, 4096}) {
    bool ok = true, any = false;
    for (auto& L : b->layers)
      for (auto& g : L.groups) {
        ok = ok && ntt_valid_split(g.params->host.N, n1);
        any = true;
      }
    if (ok && any) splits.push_back(n1);
  }
  if (splits.size() < 2) splits.assign(1, 0);
  std::vector<aoa_batch_options> cand;
  for (int n1 : splits)
    for (int xcd = 0; xcd < 2; ++xcd) {
      cand.push_back({n1, 0, xcd, -1});  // item order: no strict split
      for (int ss = -1; ss <= 1; ++ss) cand.push_back({n1, 1, xcd, ss});
    }
  std::vector<std::pair<const uint64_t*, size_t>> outs;
  output_ranges(b, &outs);
  uint64_t* acc = nullptr;
  if (hipMalloc(&acc, sizeof(uint64_t)) != hipSuccess) return -set_error(AOA_ERR_HIP, "hipMalloc failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const aoa_batch_options saved = b->opt;
  auto checksum = [&](uint64_t* v) -> int {
    if (hipMemsetAsync(acc, 0, sizeof(uint64_t), st) != hipSuccess) return set_error(AOA_ERR_HIP, "hipMemsetAsync failed");
    uint64_t off = 0;
    for (auto& r : outs) {
      launch_checksum(r.first, r.second, off, acc, st);
      off += r.second;
    }
    if (hipMemcpyAsync(v, acc, sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return set_error(AOA_ERR_HIP, "checksum copy failed");
    return check_launch("aoa_batch_autotune");
  };
  int best = -1, count = 0;
  double best_ms = 1e30;
  uint64_t base_sum = 0;
  bool reproducible = true;
  for (size_t c = 0; c < cand.size() && !rc; ++c) {
    rc = aoa_batch_set_options(b, &cand[c]);
    if (rc) break;
    rc = aoa_batch_launch(b, st, 1);  // warm-up (captures the graph)
    if (rc) break;
    if (hipStreamSynchronize(st) != hipSuccess) {
      rc = set_error(AOA_ERR_HIP, "hipStreamSynchronize failed");
      break;
    }
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && !rc; ++r) rc = aoa_batch_launch(b, st, 1);
    (void)hipEventRecord(e1, st);
    if (rc || hipEventSynchronize(e1) != hipSuccess) {
      if (!rc) rc = set_error(AOA_ERR_HIP, "hipEventSynchronize failed");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    uint64_t sum = 0;
    if ((rc = checksum(&sum))) break;
    if (c == 0) {
      // outputs that are a function of unchanged inputs repeat exactly
      base_sum = sum;
      if ((rc = aoa_batch_launch(b, st, 1))) break;
      uint64_t again = 0;
      if ((rc = checksum(&again))) break;
      reproducible = again == base_sum;
    }
    const bool exact = reproducible && sum == base_sum;
    const double t = ms / reps;
    if (count < max && results) results[count] = aoa_tune_result{cand[c], t, sum, exact ? 1 : 0};
    ++count;
    if ((exact || !reproducible) && t < best_ms) {
      best_ms = t;
      best = (int)c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(acc);
  if (rc) {
    (void)aoa_batch_set_options(b, &saved);
    return -rc;
  }
  rc = aoa_batch_set_options(b, best >= 0 ? &cand[best] : &saved);
  return rc ? -rc : count;
}
