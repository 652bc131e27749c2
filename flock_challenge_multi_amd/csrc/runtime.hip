// Host runtime of the device library: status/latch handling (CF/gpu.rs:13-21,
// 455-516), the per-device twiddle cache (the device counterpart of
// cached_standard_twiddles, CF/ntt/additive_ntt_f128.rs:146-178) and set-up
// arithmetic.
#include <hip/hip_runtime.h>
#include <immintrin.h>

#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <unordered_map>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace flk {

static std::atomic<int> g_latched{0};
static std::atomic<int> g_enabled{1};
static thread_local char g_err[512];

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int device_failure(hipError_t e, const char* where) {
  g_latched.store(1);
  return set_error(FLK_E_DEVICE, "%s: %s", where, hipGetErrorString(e));
}

static bool env_disabled() {
  const char* v = getenv("FLOCK_NO_GPU");
  return v && v[0] && strcmp(v, "0") != 0;
}

// Failure injection for tests of the latch (flk_inject_device_failures): after
// g_inject_skip more entry checks, the next g_inject_fail ones fail as a device
// runtime failure would.
static std::atomic<bool> g_inject_armed{false};
static std::mutex g_inject_mu;
static unsigned g_inject_skip = 0, g_inject_fail = 0;

static int enabled_state() {
  if (g_latched.load()) return set_error(FLK_E_DISABLED, "device path latched off after a runtime failure");
  if (!g_enabled.load()) return set_error(FLK_E_DISABLED, "device path disabled by flk_set_enabled(0)");
  if (env_disabled()) return set_error(FLK_E_DISABLED, "device path disabled by FLOCK_NO_GPU");
  return FLK_OK;
}

int check_enabled() {
  if (int rc = enabled_state()) return rc;
  if (g_inject_armed.load(std::memory_order_relaxed)) {
    std::lock_guard<std::mutex> lk(g_inject_mu);
    if (g_inject_skip) {
      --g_inject_skip;
    } else if (g_inject_fail) {
      if (--g_inject_fail == 0) g_inject_armed.store(false);
      return device_failure(hipErrorUnknown, "injected device failure");
    }
  }
  return FLK_OK;
}

// ---- host GF(2^128) for set-up constants -------------------------------
// 64x64 carry-less product: PCLMULQDQ when the host CPU has it (checked once),
// otherwise the bit-serial loop.
__attribute__((target("pclmul,sse4.1"))) static inline void clmul64_pclmul(uint64_t a, uint64_t b, uint64_t& lo,
                                                                          uint64_t& hi) {
  const __m128i p = _mm_clmulepi64_si128(_mm_set_epi64x(0, (long long)a), _mm_set_epi64x(0, (long long)b), 0x00);
  lo = (uint64_t)_mm_cvtsi128_si64(p);
  hi = (uint64_t)_mm_extract_epi64(p, 1);
}

static inline void clmul64_portable(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  uint64_t l = 0, h = 0;
  for (int i = 0; i < 64; ++i)
    if ((a >> i) & 1) {
      l ^= b << i;
      if (i) h ^= b >> (64 - i);
    }
  lo = l;
  hi = h;
}

static const bool g_has_pclmul = __builtin_cpu_supports("pclmul");

static inline void clmul64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  if (g_has_pclmul)
    clmul64_pclmul(a, b, lo, hi);
  else
    clmul64_portable(a, b, lo, hi);
}

flk_f128 host_mul(flk_f128 a, flk_f128 b) {
  uint64_t r[4], x0, x1;
  clmul64(a.lo, b.lo, r[0], r[1]);
  clmul64(a.hi, b.hi, r[2], r[3]);
  clmul64(a.lo ^ a.hi, b.lo ^ b.hi, x0, x1);  // Karatsuba middle term
  x0 ^= r[0] ^ r[2];
  x1 ^= r[1] ^ r[3];
  r[1] ^= x0;
  r[2] ^= x1;
  // fold r2:r3 * x^128 = U ^ U<<1 ^ U<<2 ^ U<<7
  uint64_t u0 = r[2], u1 = r[3];
  uint64_t t0 = u0 ^ (u0 << 1) ^ (u0 << 2) ^ (u0 << 7);
  uint64_t t1 = u1 ^ ((u1 << 1) | (u0 >> 63)) ^ ((u1 << 2) | (u0 >> 62)) ^ ((u1 << 7) | (u0 >> 57));
  uint64_t ov = (u1 >> 63) ^ (u1 >> 62) ^ (u1 >> 57);
  uint64_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
  return flk_f128{r[0] ^ t0 ^ corr, r[1] ^ t1};
}

flk_f128 host_inv(flk_f128 a) {
  // a^(2^128 - 2): exponent bits 1..127 set.
  flk_f128 result{1, 0}, base = a;
  for (int i = 0; i < 128; ++i) {
    if (i >= 1) result = host_mul(result, base);
    base = host_mul(base, base);
  }
  return result;
}

flk_f128 phi8_host(uint8_t v) {
  const flk_f128 g{0x6b8330483c2e9849ull, 0x0dcb364640a222feull};  // phi8(0x02), CF/field/phi8.rs:28-31
  flk_f128 acc{0, 0}, p{1, 0};
  for (int i = 0; i < 8; ++i) {
    if ((v >> i) & 1) acc = flk_f128{acc.lo ^ p.lo, acc.hi ^ p.hi};
    p = host_mul(p, g);
  }
  return acc;
}

// L_i(z) = prod_{j != i} (z + s_j) / prod_{j != i} (s_i + s_j), s_j = phi8(base + j).
// The denominators depend on the nodes only and are inverted once per
// (k_skip, base); the numerators are prefix / suffix products of (z + s_j).
int lagrange_weights_nodes(size_t k_skip, flk_f128 z, size_t base, flk_f128* out) {
  if (k_skip > 7 || base + (size_t(1) << k_skip) > 256 || !out)
    return set_error(FLK_E_SHAPE, "lagrange_weights: k_skip=%zu base=%zu", k_skip, base);
  static std::mutex mu;
  static std::map<std::pair<size_t, size_t>, std::pair<std::vector<flk_f128>, std::vector<flk_f128>>> cache;
  const size_t ell = size_t(1) << k_skip;
  const std::vector<flk_f128>* nodes;
  const std::vector<flk_f128>* den_inv;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto& e = cache[{k_skip, base}];
    if (e.first.empty()) {
      e.first.resize(ell);
      e.second.resize(ell);
      for (size_t j = 0; j < ell; ++j) e.first[j] = phi8_host((uint8_t)(base + j));
      for (size_t i = 0; i < ell; ++i) {
        flk_f128 den{1, 0};
        for (size_t j = 0; j < ell; ++j)
          if (j != i) den = host_mul(den, flk_f128{e.first[i].lo ^ e.first[j].lo, e.first[i].hi ^ e.first[j].hi});
        e.second[i] = host_inv(den);
      }
    }
    nodes = &e.first;
    den_inv = &e.second;
  }
  std::vector<flk_f128> t(ell), pre(ell + 1), suf(ell + 1);
  for (size_t j = 0; j < ell; ++j) t[j] = flk_f128{z.lo ^ (*nodes)[j].lo, z.hi ^ (*nodes)[j].hi};
  pre[0] = flk_f128{1, 0};
  for (size_t j = 0; j < ell; ++j) pre[j + 1] = host_mul(pre[j], t[j]);
  suf[ell] = flk_f128{1, 0};
  for (size_t j = ell; j-- > 0;) suf[j] = host_mul(suf[j + 1], t[j]);
  for (size_t i = 0; i < ell; ++i) out[i] = host_mul(host_mul(pre[i], suf[i + 1]), (*den_inv)[i]);
  return FLK_OK;
}

void zc_fixed_challenges(flk_f128 out[7]) {
  static const uint8_t small[3] = {0xF7, 0x53, 0xB5};
  for (int i = 0; i < 3; ++i) out[i] = phi8_host(small[i]);
  for (int i = 0; i < 4; ++i) {
    const flk_f128 g{uint64_t(1) << (1u << i), 0};
    out[3 + i] = host_mul(g, host_inv(flk_f128{g.lo ^ 1u, 0}));
  }
}

// ---- twiddle tables ---------------------------------------------------------
// Row i of the normalised subspace evaluations: evals[i][k] = Ŵ_i(x^{i+k})
// (CF/ntt/additive_ntt_f128.rs:54-84); twiddle(l, b) = sum_j bit_j(b) evals[dim-l-1][1+j]
// (:2684-2693), materialised breadth-first with layer l at offset 2^l - 1 by the
// half-doubling construction of :108-130.
static std::vector<flk_f128> compute_twiddles(size_t dim) {
  std::vector<std::vector<flk_f128>> ev(dim);
  for (size_t k = 0; k < dim; ++k) ev[0].push_back(flk_f128{k < 64 ? (1ull << k) : 0ull, k >= 64 ? (1ull << (k - 64)) : 0ull});
  for (size_t i = 1; i < dim; ++i) {
    const auto& prev = ev[i - 1];
    for (size_t k = 1; k < prev.size(); ++k) {
      flk_f128 s{prev[k].lo ^ prev[0].lo, prev[k].hi ^ prev[0].hi};
      ev[i].push_back(host_mul(prev[k], s));
    }
  }
  for (auto& row : ev) {
    flk_f128 inv = host_inv(row[0]);
    for (auto& v : row) v = host_mul(v, inv);
  }
  std::vector<flk_f128> tw;
  tw.reserve((size_t(1) << dim) - 1);
  for (size_t layer = 0; layer < dim; ++layer) {
    size_t start = tw.size();
    const auto& row = ev[dim - layer - 1];
    tw.push_back(flk_f128{0, 0});
    for (size_t bit = 0; bit + 1 < row.size(); ++bit) {
      size_t half = size_t(1) << bit;
      for (size_t b = 0; b < half; ++b) {
        flk_f128 v = tw[start + b];
        tw.push_back(flk_f128{v.lo ^ row[1 + bit].lo, v.hi ^ row[1 + bit].hi});
      }
    }
  }
  return tw;
}

static std::mutex g_tw_mu;
static std::map<size_t, std::vector<flk_f128>> g_tw_host;
static std::map<std::pair<int, size_t>, uint4*> g_tw_dev;

const flk_f128* twiddles_host(size_t dim) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw_host.find(dim);
  if (it == g_tw_host.end()) it = g_tw_host.emplace(dim, compute_twiddles(dim)).first;
  return it->second.data();
}

const uint4* twiddles_device(size_t dim, int* status) {
  *status = FLK_OK;
  if (dim == 0) return nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *status = device_failure(e, "hipGetDevice");
    return nullptr;
  }
  const flk_f128* host = twiddles_host(dim);
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, dim);
  auto it = g_tw_dev.find(key);
  if (it != g_tw_dev.end()) return it->second;
  size_t bytes = ((size_t(1) << dim) - 1) * sizeof(flk_f128);
  uint4* d = nullptr;
  e = hipMalloc(&d, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *status = set_error(FLK_E_NOMEM, "hipMalloc(%zu) for twiddles: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  e = hipMemcpy(d, host, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    *status = device_failure(e, "hipMemcpy twiddles");
    return nullptr;
  }
  g_tw_dev.emplace(key, d);
  return d;
}

// out[k] = XOR of parts[k * grid .. k * grid + grid), one work-group of 1024
// lanes per k; each lane's loads are independent (issued back to back).
constexpr int PSUM_WG = 1024;
__global__ void __launch_bounds__(PSUM_WG) partials_sum_kernel(const uint4* __restrict__ parts, uint32_t grid,
                                                               uint4* __restrict__ out) {
  __shared__ uint4 red[PSUM_WG / 64];
  const uint4* p = parts + (size_t)blockIdx.x * grid;
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll 8
  for (uint32_t i = threadIdx.x; i < grid; i += PSUM_WG) {
    const uint4 v = p[i];
    acc = make_uint4(acc.x ^ v.x, acc.y ^ v.y, acc.z ^ v.z, acc.w ^ v.w);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    acc = make_uint4(acc.x ^ __shfl_xor(acc.x, m), acc.y ^ __shfl_xor(acc.y, m), acc.z ^ __shfl_xor(acc.z, m),
                     acc.w ^ __shfl_xor(acc.w, m));
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 r = red[0];
    for (int w = 1; w < PSUM_WG / 64; ++w) r = make_uint4(r.x ^ red[w].x, r.y ^ red[w].y, r.z ^ red[w].z, r.w ^ red[w].w);
    out[blockIdx.x] = r;
  }
}


// ---------------------------------------------------------------------------
// Device arena: one hipMalloc'd block, best-fit free list with coalescing.
// Host bookkeeping only; blocks are 256-byte aligned.
// ---------------------------------------------------------------------------
class Arena {
 public:
  static constexpr size_t ALIGN = 256;
  bool init(size_t cap) {
    cap = (cap + ALIGN - 1) & ~(ALIGN - 1);
    if (hipMalloc(reinterpret_cast<void**>(&base_), cap) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    cap_ = cap;
    add_free(0, cap);
    return true;
  }
  // A view over `cap` bytes carved from `parent` (returned to it on deletion).
  bool init_view(Arena* parent, size_t cap) {
    cap = (cap + ALIGN - 1) & ~(ALIGN - 1);
    void* p = parent->alloc(cap);
    if (!p) return false;
    parent_ = parent;
    base_ = static_cast<char*>(p);
    cap_ = cap;
    add_free(0, cap);
    return true;
  }
  ~Arena() {
    if (parent_)
      parent_->release(base_);
    else if (base_)
      (void)hipFree(base_);
  }
  bool contains(const void* p) const {
    const char* c = static_cast<const char*>(p);
    return base_ && c >= base_ && c < base_ + cap_;
  }
  void* alloc(size_t bytes) {
    bytes = bytes ? (bytes + ALIGN - 1) & ~(ALIGN - 1) : ALIGN;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = by_size_.lower_bound({bytes, 0});
    if (it == by_size_.end()) {
      ++fallbacks_;
      return nullptr;
    }
    const size_t off = it->second, len = it->first;
    by_size_.erase(it);
    by_off_.erase(off);
    if (len > bytes) add_free(off + bytes, len - bytes);
    used_[off] = bytes;
    in_use_ += bytes;
    peak_ = std::max(peak_, in_use_);
    return base_ + off;
  }
  bool release(void* p) {
    if (!contains(p)) return false;
    std::lock_guard<std::mutex> lk(mu_);
    size_t off = static_cast<size_t>(static_cast<char*>(p) - base_);
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    size_t len = u->second;
    used_.erase(u);
    in_use_ -= len;
    auto next = by_off_.find(off + len);
    if (next != by_off_.end()) {
      by_size_.erase({next->second, next->first});
      len += next->second;
      by_off_.erase(next);
    }
    auto prev = by_off_.lower_bound(off);
    if (prev != by_off_.begin()) {
      --prev;
      if (prev->first + prev->second == off) {
        by_size_.erase({prev->second, prev->first});
        off = prev->first;
        len += prev->second;
        by_off_.erase(prev);
      }
    }
    add_free(off, len);
    return true;
  }
  void stats(size_t* capacity, size_t* in_use, size_t* peak, size_t* fallbacks) const {
    std::lock_guard<std::mutex> lk(mu_);
    if (capacity) *capacity = cap_;
    if (in_use) *in_use = in_use_;
    if (peak) *peak = peak_;
    if (fallbacks) *fallbacks = fallbacks_;
  }

 private:
  void add_free(size_t off, size_t len) {
    by_off_[off] = len;
    by_size_.insert({len, off});
  }
  Arena* parent_ = nullptr;
  char* base_ = nullptr;
  size_t cap_ = 0, in_use_ = 0, peak_ = 0, fallbacks_ = 0;
  std::map<size_t, size_t> by_off_;              // free: offset -> length
  std::set<std::pair<size_t, size_t>> by_size_;  // free: (length, offset)
  std::unordered_map<size_t, size_t> used_;      // in use: offset -> length
  mutable std::mutex mu_;
};

static thread_local Arena* t_arena = nullptr;
static std::mutex g_arenas_mu;
static std::vector<Arena*> g_arenas;  // live arenas (stream_free looks blocks up here)

ArenaScope::ArenaScope(Arena* a) : prev_(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev_; }

Arena* arena_new(size_t bytes, int* status) {
  *status = FLK_OK;
  Arena* a = new (std::nothrow) Arena();
  if (!a || !a->init(bytes)) {
    delete a;
    *status = set_error(FLK_E_NOMEM, "arena: hipMalloc of %zu bytes failed", bytes);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_arenas_mu);
  g_arenas.push_back(a);
  return a;
}

Arena* arena_view_new(Arena* parent, size_t bytes, int* status) {
  *status = FLK_OK;
  Arena* a = new (std::nothrow) Arena();
  if (!a || !a->init_view(parent, bytes)) {
    delete a;
    *status = set_error(FLK_E_NOMEM, "arena view: %zu bytes do not fit in the parent arena", bytes);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_arenas_mu);
  g_arenas.push_back(a);
  return a;
}

void arena_delete(Arena* a) {
  if (!a) return;
  {
    std::lock_guard<std::mutex> lk(g_arenas_mu);
    g_arenas.erase(std::remove(g_arenas.begin(), g_arenas.end(), a), g_arenas.end());
  }
  delete a;
}

void arena_stats(const Arena* a, size_t* capacity, size_t* in_use, size_t* peak, size_t* fallbacks) {
  a->stats(capacity, in_use, peak, fallbacks);
}

static void keep_pool_memory();

hipError_t stream_alloc(void** p, size_t bytes, hipStream_t st) {
  if (t_arena && (*p = t_arena->alloc(bytes)) != nullptr) return hipSuccess;
  keep_pool_memory();
  return hipMallocAsync(p, bytes, st);
}

hipError_t stream_free(void* p, hipStream_t st) {
  if (!p) return hipSuccess;
  if (t_arena && t_arena->release(p)) return hipSuccess;
  {
    // Newest first: a view is created after the arena it is carved from, so a
    // block of a view is released there, never as the view's block of the parent.
    std::lock_guard<std::mutex> lk(g_arenas_mu);
    bool inside = false;
    for (auto it = g_arenas.rbegin(); it != g_arenas.rend(); ++it)
      if ((*it)->contains(p)) {
        inside = true;
        if ((*it)->release(p)) return hipSuccess;
      }
    if (inside) return hipErrorInvalidValue;
  }
  return hipFreeAsync(p, st);
}

// The default stream-ordered pool of each device keeps what the reduction
// scratch frees (release threshold raised once per device), so repeated calls
// reuse it instead of returning it to the HIP runtime at every synchronisation.
static void keep_pool_memory() {
  static std::mutex mu;
  static bool done[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (done[dev]) return;
  done[dev] = true;
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  (void)hipGetLastError();
}

__global__ void zero_n_kernel(uint4* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Small device -> host reads (<= 64 KiB, the round messages and challenges of
// the provers) go through the calling thread's mailbox: coherent, mapped pinned
// memory that a one-work-group kernel fills with system-scope stores before it
// sets a sequence word; the host spins on that word instead of synchronizing
// the stream (7 us against 14 us per read-back on the MI355X box).  When the
// word has not appeared after MAILBOX_SPIN (a long kernel ahead on the stream,
// or a failed one) the host synchronizes the stream, which also reports a
// device failure.  Seeing the word means every earlier operation on the stream
// has completed, as after a synchronization.
constexpr uint32_t MAILBOX_BYTES = 64u << 10;
constexpr double MAILBOX_SPIN_S = 2e-3;

__global__ void __launch_bounds__(256) publish_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ box,
                                                      uint32_t bytes, uint32_t* flag, uint32_t seq) {
  if (((uintptr_t)src | bytes) % 16 == 0) {
    for (uint32_t i = threadIdx.x; i < bytes / 16; i += blockDim.x)
      reinterpret_cast<uint4*>(box)[i] = reinterpret_cast<const uint4*>(src)[i];
  } else {
    for (uint32_t i = threadIdx.x; i < bytes; i += blockDim.x) box[i] = src[i];
  }
  // every thread's stores, ordered before thread 0's fence by the barrier, reach
  // host memory before the flag (one system-scope fence, not one per thread)
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Host -> device uploads staged through a per-thread pinned ring (8 MiB,
// allocated on first use): the bytes are copied into the ring and the DMA is
// enqueued on the stream, so the caller's (pageable) buffer may be reused at
// once and no synchronisation is needed.  A region of the ring is reused only
// after the stream that read it has been seen to pass its copies: every
// read-back through read_to_host on that stream waits for the whole stream and
// recycles the ring; a full ring, or an upload on another stream, synchronises
// the stream the pending copies belong to first.  Larger uploads are copied
// and waited for.
struct UploadRing {
  uint8_t* base = nullptr;
  size_t off = 0;
  hipStream_t pending = nullptr;  // the stream of the copies in [0, off)
  bool failed = false;
};
constexpr size_t UPLOAD_RING_BYTES = size_t(8) << 20;
static thread_local UploadRing t_ring;

static hipError_t ring_drain() {
  UploadRing& r = t_ring;
  if (r.off == 0) return hipSuccess;
  hipError_t e = hipStreamSynchronize(r.pending);
  r.off = 0;
  r.pending = nullptr;
  return e;
}

hipError_t upload_staged(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  UploadRing& r = t_ring;
  if (!r.base && !r.failed) {
    void* h = nullptr;
    if (hipHostMalloc(&h, UPLOAD_RING_BYTES, hipHostMallocDefault) == hipSuccess) {
      r.base = static_cast<uint8_t*>(h);
    } else {
      (void)hipGetLastError();
      r.failed = true;
    }
  }
  const size_t need = (bytes + 255) & ~size_t(255);
  if (!r.base || need > UPLOAD_RING_BYTES) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
  }
  if (r.off && (r.pending != st || r.off + need > UPLOAD_RING_BYTES)) {
    hipError_t e = ring_drain();
    if (e != hipSuccess) return e;
  }
  uint8_t* slot = r.base + r.off;
  memcpy(slot, src, bytes);
  hipError_t e = hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  r.off += need;
  r.pending = st;
  return hipSuccess;
}

// The stream st has completed everything enqueued before this point: the ring's
// copies on it are done.
static void ring_passed(hipStream_t st) {
  UploadRing& r = t_ring;
  if (r.off && r.pending == st) {
    r.off = 0;
    r.pending = nullptr;
  }
}

struct Mailbox {
  uint8_t* host = nullptr;  // MAILBOX_BYTES of data, then the sequence word
  uint8_t* dev = nullptr;
  uint32_t seq = 0;
  int device = -1;
  bool failed = false;
};

// Synchronous device -> host read through a per-thread pinned bounce buffer
// (1 MiB, allocated on first use and kept): a pageable destination makes the
// runtime stage the copy itself, 23 us per small read-back on the MI355X box
// against 14 us through pinned memory.
static hipError_t read_bounce(void* dst, const void* src, size_t bytes, hipStream_t st) {
  constexpr size_t BOUNCE = size_t(1) << 20;
  static thread_local void* bounce = nullptr;
  if (bytes <= BOUNCE && !bounce) {
    if (hipHostMalloc(&bounce, BOUNCE, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      bounce = nullptr;
    }
  }
  hipError_t e;
  if (bytes <= BOUNCE && bounce) {
    e = hipMemcpyAsync(bounce, src, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) memcpy(dst, bounce, bytes);
  } else {
    e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  if (e == hipSuccess) ring_passed(st);
  return e;
}

// The calling thread's mailbox on the current device, made on first use (one
// per thread, for its first device; NULL when unavailable: a thread that moved
// to another device, or no mapped pinned memory).
static thread_local Mailbox t_mb;
static Mailbox* mailbox_current() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  Mailbox& mb = t_mb;
  if (mb.device == -1 && !mb.failed) {
    void* h = nullptr;
    void* d = nullptr;
    if (hipHostMalloc(&h, MAILBOX_BYTES + 256, hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess &&
        hipHostGetDevicePointer(&d, h, 0) == hipSuccess) {
      memset(h, 0, MAILBOX_BYTES + 256);
      mb.host = static_cast<uint8_t*>(h);
      mb.dev = static_cast<uint8_t*>(d);
      mb.device = dev;
    } else {
      (void)hipGetLastError();
      if (h) (void)hipHostFree(h);
      mb.failed = true;
    }
  }
  return (mb.failed || mb.device != dev) ? nullptr : &mb;
}

// Waits until the mailbox's sequence word reads seq (the stream is past the
// store), synchronising the stream after MAILBOX_SPIN_S, then copies the bytes.
// An armed reduction sets the word from inside a kernel that may be followed by
// copies enqueued since, so only a publish kernel's word (or a synchronisation)
// lets the upload ring recycle.
// stream_done: the word was set by the last operation enqueued on st (the publish
// kernel), so everything before it, the upload ring's copies included, is done.
static hipError_t mailbox_wait(Mailbox& mb, uint32_t seq, void* dst, size_t bytes, hipStream_t st,
                               bool stream_done) {
  const uint32_t* flag_host = reinterpret_cast<const uint32_t*>(mb.host + MAILBOX_BYTES);
  const auto t0 = std::chrono::steady_clock::now();
  bool seen = false;
  for (;;) {
    if (__atomic_load_n(flag_host, __ATOMIC_ACQUIRE) == seq) {
      seen = true;
      break;
    }
    if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > MAILBOX_SPIN_S) break;
    _mm_pause();
  }
  if (!seen) {
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (__atomic_load_n(flag_host, __ATOMIC_ACQUIRE) != seq) return hipErrorUnknown;
  }
  memcpy(dst, mb.host, bytes);
  if (stream_done || !seen) ring_passed(st);
  return hipSuccess;
}

hipError_t read_to_host(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  Mailbox* mb = bytes <= MAILBOX_BYTES ? mailbox_current() : nullptr;
  if (!mb) return read_bounce(dst, src, bytes, st);
  const uint32_t seq = ++mb->seq;
  uint32_t* flag_dev = reinterpret_cast<uint32_t*>(mb->dev + MAILBOX_BYTES);
  hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(256), 0, st, static_cast<const uint8_t*>(src), mb->dev,
                     (uint32_t)bytes, flag_dev, seq);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return mailbox_wait(*mb, seq, dst, bytes, st, true);
}

// ---------------------------------------------------------------------------
// Armed read-backs.  flk_read_arm(dst, bytes) hands the next reduction on this
// thread whose result is `bytes` long the calling thread's mailbox: the launch
// that sums its work-group partials (partials_sum_publish_kernel) also stores the
// result there and sets the sequence word, so no publish kernel runs after it;
// flk_read_armed then waits for the word and copies the bytes to dst.  When no
// reduction took the arm, flk_read_armed reads the result the ordinary way.
// ---------------------------------------------------------------------------
struct ReadArm {
  void* dst = nullptr;
  size_t bytes = 0;
  bool armed = false, taken = false;
  uint32_t seq = 0;
};
static thread_local ReadArm t_arm;

static bool take_read_arm(size_t bytes, uint4** box, uint32_t** flag, uint32_t* seq) {
  ReadArm& a = t_arm;
  if (!a.armed || a.taken || a.bytes != bytes) return false;
  Mailbox* mb = mailbox_current();
  if (!mb) return false;
  a.taken = true;
  a.seq = *seq = ++mb->seq;
  *box = reinterpret_cast<uint4*>(mb->dev);
  *flag = reinterpret_cast<uint32_t*>(mb->dev + MAILBOX_BYTES);
  return true;
}

// out[k] = box[k] = XOR of parts[k * grid .. k * grid + grid) for k < n_out, one
// work-group of 1024 lanes, then the mailbox's sequence word (after a
// system-scope fence, as publish_kernel).
__global__ void __launch_bounds__(1024) partials_sum_publish_kernel(const uint4* __restrict__ parts, uint32_t grid,
                                                                   uint32_t n_out, uint4* __restrict__ out,
                                                                   uint4* __restrict__ box, uint32_t* flag,
                                                                   uint32_t seq) {
  __shared__ uint4 red[4][16];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t k = 0; k < n_out; ++k) {
    const uint4* p = parts + (size_t)k * grid;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll 8
    for (uint32_t i = threadIdx.x; i < grid; i += 1024) {
      const uint4 v = p[i];
      acc = make_uint4(acc.x ^ v.x, acc.y ^ v.y, acc.z ^ v.z, acc.w ^ v.w);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
      acc = make_uint4(acc.x ^ __shfl_xor(acc.x, m), acc.y ^ __shfl_xor(acc.y, m), acc.z ^ __shfl_xor(acc.z, m),
                       acc.w ^ __shfl_xor(acc.w, m));
    if (lane == 0) red[k][wave] = acc;
  }
  __syncthreads();
  if (threadIdx.x < n_out) {
    uint4 r = red[threadIdx.x][0];
    for (uint32_t w = 1; w < 16; ++w) {
      const uint4 v = red[threadIdx.x][w];
      r = make_uint4(r.x ^ v.x, r.y ^ v.y, r.z ^ v.z, r.w ^ v.w);
    }
    out[threadIdx.x] = r;
    box[threadIdx.x] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();  // the results reach host memory before the sequence word
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

int reduce_begin(unsigned grid, int n_out, uint4* out, hipStream_t st, ReduceTail* rt) {
  rt->out = out;
  rt->box = nullptr;
  rt->flag = nullptr;
  rt->seq = 0;
  int rc = partials_alloc(&rt->parts, grid, n_out, out, st);
  if (rc) return rc;
  if (n_out <= 4) take_read_arm((size_t)n_out * sizeof(uint4), &rt->box, &rt->flag, &rt->seq);
  return FLK_OK;
}

int reduce_end(const ReduceTail& rt, unsigned grid, int n_out, hipStream_t st) {
  if (!rt.box) return partials_sum_free(rt.parts, grid, n_out, rt.out, st);
  // with grid == 1 the partials are out itself: the launch only publishes them
  hipLaunchKernelGGL(partials_sum_publish_kernel, dim3(1), dim3(1024), 0, st, rt.parts, grid, (uint32_t)n_out, rt.out,
                     rt.box, rt.flag, rt.seq);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return device_failure(e, "partials_sum_publish_kernel");
  if (rt.parts == rt.out) return FLK_OK;
  e = stream_free(rt.parts, st);
  if (e != hipSuccess) return device_failure(e, "reduction scratch: free");
  return FLK_OK;
}

int zero_n(uint4* p, uint64_t n, hipStream_t st) {
  if (!n) return FLK_OK;
  const uint64_t wgs = (n + 255) / 256;
  hipLaunchKernelGGL(zero_n_kernel, dim3((unsigned)(wgs < 65535 ? wgs : 65535)), dim3(256), 0, st, p, n);
  FLK_LAUNCH_CHECK("zero_n_kernel");
  return FLK_OK;
}

int partials_alloc(uint4** parts, unsigned grid, int n_out, uint4* out, hipStream_t st) {
  *parts = nullptr;
  if (grid == 1) {  // parts[k * 1 + 0] is out[k]: no scratch, no second launch
    *parts = out;
    return FLK_OK;
  }
  hipError_t e = stream_alloc(reinterpret_cast<void**>(parts), (size_t)grid * (size_t)n_out * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *parts = nullptr;
    return set_error(FLK_E_NOMEM, "reduction scratch: allocation of %zu bytes failed",
                     (size_t)grid * (size_t)n_out * sizeof(uint4));
  }
  return FLK_OK;
}

int partials_sum(const uint4* parts, unsigned grid, int n_out, uint4* out, hipStream_t st) {
  hipLaunchKernelGGL(partials_sum_kernel, dim3((unsigned)n_out), dim3(PSUM_WG), 0, st, parts, (uint32_t)grid, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return device_failure(e, "partials_sum_kernel");
  return FLK_OK;
}

int partials_sum_free(uint4* parts, unsigned grid, int n_out, uint4* out, hipStream_t st) {
  if (parts == out) return FLK_OK;  // one work-group wrote out directly
  hipLaunchKernelGGL(partials_sum_kernel, dim3((unsigned)n_out), dim3(PSUM_WG), 0, st, parts, (uint32_t)grid, out);
  hipError_t e = hipGetLastError();
  const hipError_t ef = stream_free(parts, st);
  if (e != hipSuccess) return device_failure(e, "partials_sum_kernel");
  if (ef != hipSuccess) return device_failure(ef, "reduction scratch: free");
  return FLK_OK;
}

}  // namespace flk

extern "C" {

int flk_device_available(void) {
  if (flk::check_enabled() != FLK_OK) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n > 0;
}

int flk_is_disabled(void) { return flk::enabled_state() != FLK_OK; }

void flk_undisable(void) { flk::g_latched.store(0); }

void flk_inject_device_failures(unsigned skip, unsigned n) {
  std::lock_guard<std::mutex> lk(flk::g_inject_mu);
  flk::g_inject_skip = n ? skip : 0;
  flk::g_inject_fail = n;
  flk::g_inject_armed.store(n != 0);
}

void flk_set_enabled(int enabled) { flk::g_enabled.store(enabled ? 1 : 0); }

const char* flk_last_error(void) { return flk::g_err; }

const char* flk_version(void) { return "flock-mi355x 0.1 (gfx950)"; }

flk_f128 flk_f128_mul(flk_f128 a, flk_f128 b) { return flk::host_mul(a, b); }

int flk_alloc(size_t bytes, void** out) {
  if (!out) return flk::set_error(FLK_E_SHAPE, "flk_alloc: null output");
  *out = nullptr;
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (bytes == 0) return FLK_OK;
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return flk::set_error(FLK_E_NOMEM, "flk_alloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return FLK_OK;
}

int flk_free(void* p) {
  if (!p) return FLK_OK;
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return flk::device_failure(e, "flk_free");
  return FLK_OK;
}

int flk_alloc_async(size_t bytes, void* stream, void** out) {
  if (!out) return flk::set_error(FLK_E_SHAPE, "flk_alloc_async: null output");
  *out = nullptr;
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (bytes == 0) return FLK_OK;
  hipError_t e = flk::stream_alloc(out, bytes, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return flk::set_error(FLK_E_NOMEM, "flk_alloc_async(%zu): %s", bytes, hipGetErrorString(e));
  }
  return FLK_OK;
}

int flk_free_async(void* p, void* stream) {
  if (!p) return FLK_OK;
  hipError_t e = flk::stream_free(p, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return flk::device_failure(e, "flk_free_async");
  return FLK_OK;
}

int flk_copy_to_device(void* dst, const void* src, size_t bytes, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!bytes) return FLK_OK;
  if (!dst || !src) return flk::set_error(FLK_E_SHAPE, "flk_copy_to_device: null pointer");
  hipError_t e = flk::upload_staged(dst, src, bytes, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return flk::device_failure(e, "flk_copy_to_device");
  return FLK_OK;
}

int flk_copy_to_host(void* dst, const void* src, size_t bytes, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!bytes) return FLK_OK;
  if (!dst || !src) return flk::set_error(FLK_E_SHAPE, "flk_copy_to_host: null pointer");
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return flk::device_failure(e, "flk_copy_to_host");
  return FLK_OK;
}

int flk_read_to_host(void* dst, const void* src, size_t bytes, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!bytes) return FLK_OK;
  if (!dst || !src) return flk::set_error(FLK_E_SHAPE, "flk_read_to_host: null pointer");
  hipError_t e = flk::read_to_host(dst, src, bytes, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return flk::device_failure(e, "flk_read_to_host");
  return FLK_OK;
}

int flk_read_arm(void* dst, size_t bytes) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!dst || !bytes || bytes > flk::MAILBOX_BYTES)
    return flk::set_error(FLK_E_SHAPE, "flk_read_arm: dst must be non-null and 1 <= bytes <= %u", flk::MAILBOX_BYTES);
  flk::t_arm = flk::ReadArm{};
  flk::t_arm.dst = dst;
  flk::t_arm.bytes = bytes;
  flk::t_arm.armed = true;
  return FLK_OK;
}

int flk_read_armed(const void* src, size_t bytes, void* stream) {
  const flk::ReadArm a = flk::t_arm;
  flk::t_arm = flk::ReadArm{};
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!a.armed || a.bytes != bytes || !src)
    return flk::set_error(FLK_E_SHAPE, "flk_read_armed: no read-back of %zu bytes armed on this thread", bytes);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (a.taken)
    e = flk::mailbox_wait(flk::t_mb, a.seq, a.dst, bytes, st, false);
  else
    e = flk::read_to_host(a.dst, src, bytes, st);
  if (e != hipSuccess) return flk::device_failure(e, "flk_read_armed");
  return FLK_OK;
}

int flk_copy_device_to_device(void* dst, const void* src, size_t bytes, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!bytes) return FLK_OK;
  if (!dst || !src) return flk::set_error(FLK_E_SHAPE, "flk_copy_device_to_device: null pointer");
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return flk::device_failure(e, "flk_copy_device_to_device");
  return FLK_OK;
}

int flk_synchronize(void* stream) {
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return flk::device_failure(e, "flk_synchronize");
  flk::ring_passed(static_cast<hipStream_t>(stream));
  return FLK_OK;
}

int flk_zc_lagrange_weights(size_t k_skip, flk_f128 z, flk_f128* out) {
  // lagrange_weights_naive (CF/zerocheck/multilinear.rs:285-330) on the skip
  // domain S = phi8(0 .. 2^k_skip - 1).
  if (k_skip > 7 || !out) return flk::set_error(FLK_E_SHAPE, "zc_lagrange_weights: k_skip=%zu", k_skip);
  return flk::lagrange_weights_nodes(k_skip, z, 0, out);
}

int flk_ntt_twiddles(size_t dim, flk_f128* out) {
  if (dim > 30) return flk::set_error(FLK_E_SHAPE, "flk_ntt_twiddles: dim %zu too large", dim);
  if (!out && dim) return flk::set_error(FLK_E_SHAPE, "flk_ntt_twiddles: null output");
  if (dim == 0) return FLK_OK;
  const flk_f128* t = flk::twiddles_host(dim);
  memcpy(out, t, ((size_t(1) << dim) - 1) * sizeof(flk_f128));
  return FLK_OK;
}

}  // extern "C"
