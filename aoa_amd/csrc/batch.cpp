// Batched (stage-fused) evaluation: many ciphertexts, mixed levels and ring
// dimensions, one launch per pipeline stage per parameter set, the whole
// sequence captured once as a hipGraph and replayed.
//
// A batch is a list of layers; a layer holds independent operations
// (HomMul+relin+rescale or rotate[-and-accumulate]) grouped by parameter set.
// Layers run in order (each may consume the previous one's outputs, as the
// rotate-and-sum steps of ckks_sumslots do, aoa-ntt/src/ckks/ckks.cpp:479-495).
// Every stage of a group is one descriptor-driven kernel launch covering all
// its ciphertexts, so a launch carries thousands of workgroups even when each
// ciphertext is small, and evaluation keys are shared between them.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "internal.hpp"

using namespace aoa;

namespace {

struct MulOp {
  aoa_mul_item it;
};
struct RotOp {
  aoa_rot_item it;
};

struct Group {
  const aoa_params* params = nullptr;
  const uint64_t* evk_b = nullptr;
  const uint64_t* evk_a = nullptr;
  std::vector<aoa_mul_item> muls;
  std::vector<aoa_rot_item> rots;
  // built: descriptor tables (device) and counts
  NttDesc* d_intt3 = nullptr;  // mul: INTT of d0, d1, d2
  int n_intt3 = 0;
  TensorDesc* d_tensor = nullptr;
  int n_tensor = 0;
  AutoDesc* d_auto = nullptr;
  int n_auto = 0;
  NttDesc* d_fwd = nullptr;  // NTT of ModUp output (and of the rotated c1)
  int n_fwd = 0;
  ModUpDesc* d_modup = nullptr;
  int n_modup = 0;
  int max_level = 0;
  NttDesc* d_key = nullptr;  // key product + INTT
  int n_key = 0;
  ModDownDesc* d_moddown = nullptr;
  int n_moddown = 0;
};

struct Layer {
  std::vector<Group> groups;
};

size_t mul_ws_limbs(int L, int K) { return 6 * (size_t)L + 3 * (size_t)K; }
size_t rot_ws_limbs(int L, int K) { return 5 * (size_t)L + 3 * (size_t)K; }

}  // namespace

struct aoa_batch {
  std::vector<Layer> layers;
  bool built = false;
  void* dmem = nullptr;  // descriptors
  void* ws = nullptr;    // workspace
  size_t ws_bytes = 0;
  size_t desc_bytes = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static Group& group_for(aoa_batch* b, const aoa_params* p, const uint64_t* kb, const uint64_t* ka) {
  if (b->layers.empty()) b->layers.emplace_back();
  Layer& L = b->layers.back();
  for (auto& g : L.groups)
    if (g.params == p && g.evk_b == kb && g.evk_a == ka) return g;
  L.groups.emplace_back();
  Group& g = L.groups.back();
  g.params = p;
  g.evk_b = kb;
  g.evk_a = ka;
  return g;
}

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
  for (int i = 0; i < n; ++i) {
    const aoa_mul_item& it = items[i];
    if (!it.c1[0] || !it.c1[1] || !it.c2[0] || !it.c2[1] || !it.out[0] || !it.out[1])
      return set_error(AOA_ERR_ARG, "null pointer in item");
    if (it.level == 0) return set_error(AOA_ERR_LEVEL, "The ciphertext level is 0. No more rescaling is possible.");
    if (it.level < 0 || it.level >= p->host.k) return set_error(AOA_ERR_ARG, "level out of range");
  }
  Group& g = group_for(b, p, evk_b, evk_a);
  g.muls.insert(g.muls.end(), items, items + n);
  return AOA_OK;
}

extern "C" int aoa_batch_add_rotate(aoa_batch* b, const aoa_params* p, const aoa_rot_item* items, int n) {
  if (!b || !p || (!items && n > 0)) return set_error(AOA_ERR_ARG, "null argument");
  if (b->built) return set_error(AOA_ERR_ARG, "batch already built");
  for (int i = 0; i < n; ++i) {
    const aoa_rot_item& it = items[i];
    if (!it.in[0] || !it.in[1] || !it.out[0] || !it.out[1] || !it.rtk_b || !it.rtk_a)
      return set_error(AOA_ERR_ARG, "null pointer in item");
    if (it.level < 0 || it.level >= p->host.k || it.rot_slots < 0) return set_error(AOA_ERR_ARG, "bad item");
  }
  // rotation groups carry their keys per item
  Group& g = group_for(b, p, nullptr, nullptr);
  g.rots.insert(g.rots.end(), items, items + n);
  return AOA_OK;
}

extern "C" int aoa_batch_barrier(aoa_batch* b) {
  if (!b) return set_error(AOA_ERR_ARG, "null batch");
  if (b->built) return set_error(AOA_ERR_ARG, "batch already built");
  if (!b->layers.back().groups.empty()) b->layers.emplace_back();
  return AOA_OK;
}

// Builds every descriptor table on the host, then copies them in one go.
extern "C" int aoa_batch_build(aoa_batch* b) {
  if (!b) return set_error(AOA_ERR_ARG, "null batch");
  if (b->built) return AOA_OK;
  // workspace: per layer, sum of per-item needs; the batch keeps the max
  size_t ws_limb_bytes = 0;
  for (auto& L : b->layers) {
    size_t s = 0;
    for (auto& g : L.groups) {
      const size_t N = g.params->host.N, K = g.params->host.K;
      for (auto& it : g.muls) s += mul_ws_limbs(it.level + 1, (int)K) * N * 8;
      for (auto& it : g.rots) s += rot_ws_limbs(it.level + 1, (int)K) * N * 8;
    }
    ws_limb_bytes = std::max(ws_limb_bytes, s);
  }
  b->ws_bytes = ws_limb_bytes;
  if (b->ws_bytes && hipMalloc(&b->ws, b->ws_bytes) != hipSuccess)
    return set_error(AOA_ERR_HIP, "hipMalloc of batch workspace failed");

  // host descriptor staging, one byte vector; device pointers patched after
  std::vector<char> host;
  struct Patch {
    Group* g;
    int which;
    size_t off;
  };
  std::vector<Patch> patches;
  auto put = [&](const void* p, size_t bytes) {
    size_t off = (host.size() + 255) & ~(size_t)255;
    host.resize(off + bytes);
    std::memcpy(host.data() + off, p, bytes);
    return off;
  };
  uint64_t* wsb = (uint64_t*)b->ws;
  for (auto& L : b->layers) {
    uint64_t* w = wsb;  // layers reuse the workspace
    for (auto& g : L.groups) {
      const DevParams& P = g.params->dev;
      const size_t N = (size_t)P.N;
      const int k = P.k, K = P.K;
      std::vector<TensorDesc> tens;
      std::vector<NttDesc> intt3, fwd, key;
      std::vector<ModUpDesc> mu;
      std::vector<ModDownDesc> md;
      std::vector<AutoDesc> au;
      g.max_level = 0;
      for (auto& it : g.muls) {
        const int lv = it.level, Lq = lv + 1;
        g.max_level = std::max(g.max_level, lv);
        uint64_t* X = w;
        uint64_t* D0 = X + (size_t)(Lq + K) * N;
        uint64_t* D1 = D0 + (size_t)Lq * N;
        uint64_t* D2 = D1 + (size_t)Lq * N;
        uint64_t* P0 = D2 + (size_t)Lq * N;
        uint64_t* P1 = P0 + (size_t)(Lq + K) * N;
        w += mul_ws_limbs(Lq, K) * N;
        for (int i = 0; i < Lq; ++i) {
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
        mu.push_back({D2, X + (size_t)Lq * N, lv, 0});
        for (int j = 0; j < K; ++j) fwd.push_back({X + (size_t)(Lq + j) * N, X + (size_t)(Lq + j) * N, nullptr, k + j, 0});
        for (int y = 0; y < 2; ++y) {
          uint64_t* Py = y ? P1 : P0;
          const uint64_t* ky = y ? g.evk_a : g.evk_b;
          for (int m = 0; m < Lq + K; ++m) {
            const int pr = m < Lq ? m : k + (m - Lq);
            key.push_back({X + (size_t)m * N, Py + (size_t)m * N, ky + (size_t)pr * N, pr, 0});
          }
          md.push_back({Py, Py + (size_t)Lq * N, y ? D1 : D0, it.out[y], lv, 1});
        }
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
          AutoDesc a0 = {it.in[0] + o, R0 + o, it.accumulate ? it.in[0] + o : nullptr, ginv, i, 0};
          AutoDesc a1 = {it.in[1] + o, R1 + o, nullptr, ginv, i, 0};
          au.push_back(a0);
          au.push_back(a1);
          fwd.push_back({R1 + o, X + o, nullptr, i, 0});
        }
        mu.push_back({R1, X + (size_t)Lq * N, lv, 0});
        for (int j = 0; j < K; ++j) fwd.push_back({X + (size_t)(Lq + j) * N, X + (size_t)(Lq + j) * N, nullptr, k + j, 0});
        for (int y = 0; y < 2; ++y) {
          uint64_t* Py = y ? P1 : P0;
          const uint64_t* ky = y ? it.rtk_a : it.rtk_b;
          for (int m = 0; m < Lq + K; ++m) {
            const int pr = m < Lq ? m : k + (m - Lq);
            key.push_back({X + (size_t)m * N, Py + (size_t)m * N, ky + (size_t)pr * N, pr, 0});
          }
          const uint64_t* add = y == 0 ? R0 : (it.accumulate ? it.in[1] : nullptr);
          md.push_back({Py, Py + (size_t)Lq * N, add, it.out[y], lv, 0});
        }
      }
      g.n_tensor = (int)tens.size();
      g.n_intt3 = (int)intt3.size();
      g.n_auto = (int)au.size();
      g.n_fwd = (int)fwd.size();
      g.n_modup = (int)mu.size();
      g.n_key = (int)key.size();
      g.n_moddown = (int)md.size();
      if (!tens.empty()) patches.push_back({&g, 0, put(tens.data(), tens.size() * sizeof(TensorDesc))});
      if (!intt3.empty()) patches.push_back({&g, 1, put(intt3.data(), intt3.size() * sizeof(NttDesc))});
      if (!au.empty()) patches.push_back({&g, 2, put(au.data(), au.size() * sizeof(AutoDesc))});
      if (!fwd.empty()) patches.push_back({&g, 3, put(fwd.data(), fwd.size() * sizeof(NttDesc))});
      if (!mu.empty()) patches.push_back({&g, 4, put(mu.data(), mu.size() * sizeof(ModUpDesc))});
      if (!key.empty()) patches.push_back({&g, 5, put(key.data(), key.size() * sizeof(NttDesc))});
      if (!md.empty()) patches.push_back({&g, 6, put(md.data(), md.size() * sizeof(ModDownDesc))});
    }
  }
  b->desc_bytes = host.size();
  if (b->desc_bytes) {
    if (hipMalloc(&b->dmem, b->desc_bytes) != hipSuccess) return set_error(AOA_ERR_HIP, "hipMalloc of descriptors failed");
    if (hipMemcpy(b->dmem, host.data(), b->desc_bytes, hipMemcpyHostToDevice) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipMemcpy of descriptors failed");
  }
  char* base = (char*)b->dmem;
  for (auto& pt : patches) {
    void* p = base + pt.off;
    switch (pt.which) {
      case 0: pt.g->d_tensor = (TensorDesc*)p; break;
      case 1: pt.g->d_intt3 = (NttDesc*)p; break;
      case 2: pt.g->d_auto = (AutoDesc*)p; break;
      case 3: pt.g->d_fwd = (NttDesc*)p; break;
      case 4: pt.g->d_modup = (ModUpDesc*)p; break;
      case 5: pt.g->d_key = (NttDesc*)p; break;
      case 6: pt.g->d_moddown = (ModDownDesc*)p; break;
    }
  }
  b->built = true;
  return AOA_OK;
}

// Issues every stage launch of the batch on stream st.
static int enqueue(aoa_batch* b, hipStream_t st) {
  for (auto& L : b->layers) {
    // stage order inside a layer; groups of the same stage back to back
    for (auto& g : L.groups)
      if (g.n_tensor) launch_tensor_desc(g.params->dev, g.d_tensor, g.n_tensor, st);
    for (auto& g : L.groups)
      if (g.n_auto) launch_automorphism_desc(g.params->dev, g.d_auto, g.n_auto, st);
    for (auto& g : L.groups)
      if (g.n_intt3 && ntt_inverse_desc(g.params->dev, g.d_intt3, g.n_intt3, 0, st))
        return set_error(AOA_ERR_UNSUPPORTED, "no NTT split");
    for (auto& g : L.groups)
      if (g.n_modup) launch_modup_desc(g.params->dev, g.d_modup, g.n_modup, g.max_level, st);
    for (auto& g : L.groups)
      if (g.n_fwd && ntt_forward_desc(g.params->dev, g.d_fwd, g.n_fwd, 0, st))
        return set_error(AOA_ERR_UNSUPPORTED, "no NTT split");
    for (auto& g : L.groups)
      if (g.n_key && ntt_inverse_desc(g.params->dev, g.d_key, g.n_key, 0, st))
        return set_error(AOA_ERR_UNSUPPORTED, "no NTT split");
    for (auto& g : L.groups)
      if (g.n_moddown) launch_moddown_desc(g.params->dev, g.d_moddown, g.n_moddown, st);
  }
  return check_launch("aoa_batch");
}

extern "C" size_t aoa_batch_workspace_bytes(const aoa_batch* b) { return b ? b->ws_bytes : 0; }

extern "C" int aoa_batch_launch(aoa_batch* b, void* stream, int use_graph) {
  if (!b) return set_error(AOA_ERR_ARG, "null batch");
  int rc = aoa_batch_build(b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return enqueue(b, st);
  if (!b->exec) {
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipStreamCreate failed");
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return set_error(AOA_ERR_HIP, "hipStreamBeginCapture failed");
    int erc = enqueue(b, b->cap);
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
