// Proof-of-work grinding on the device (CF = crates/flock-core/src/), BLAKE3
// transcripts:
//
//   flk_pow_scan    pow_scan / blake3_pow_scan   CF/challenger.rs:605-640,1648-1660
//   flk_pow_grind   grind_pow's nonce search     CF/challenger.rs:384-474
//
// A nonce's hash is blake3(state_digest || u64le(nonce) || 24 zero bytes), one
// root-flagged single-block compression (blake3_pow_preimage, :549-554); it
// passes with at least `bits` leading zero bits, byte 0's most significant bit
// first (:558-571), i.e. the big-endian reading of output words 0 and 1.
//
// One lane hashes one nonce per sweep; sweep k covers nonces [base, base + T)
// for T threads, in increasing order, and a match lowers a device-scope
// minimum.  A sweep starts only while its first nonce is at most the minimum
// found so far, so every nonce below the result has been hashed: the answer
// is the smallest match, as the reference's ordered searches return.  The
// digest words are kernel arguments (scalar registers), round 1's column step
// (digest words only) runs once per lane, and the message words 10..15 are
// zero and fold away: 52 G functions per nonce, no memory traffic beyond one
// 8-byte minimum.  grind sweeps about half the expected work at a time.
#include <hip/hip_runtime.h>

#include <mutex>

#include "blake3_dev.hpp"
#include "internal.hpp"

namespace flk {
namespace {

constexpr int POW_WG = 256;
constexpr unsigned POW_MAX_GRID = 8192;  // 2^21 lanes per sweep
constexpr uint64_t NONE = ~0ull;

__global__ void pow_init_kernel(unsigned long long* best) {
  if (threadIdx.x == 0) *best = NONE;
}

__global__ __launch_bounds__(POW_WG) void pow_scan_kernel(uint4 d_lo, uint4 d_hi, uint64_t start, uint64_t end,
                                                          uint32_t bits, unsigned long long* best) {
  const uint64_t T = (uint64_t)gridDim.x * POW_WG;
  const uint64_t tid = (uint64_t)blockIdx.x * POW_WG + threadIdx.x;
  const uint64_t span = end - start;  // > 0
  const uint64_t sweeps = span / T + (span % T != 0);
  const uint32_t dig[8] = {d_lo.x, d_lo.y, d_lo.z, d_lo.w, d_hi.x, d_hi.y, d_hi.z, d_hi.w};
  // Round 1's column step reads only the digest words: once per lane.
  uint32_t c0 = 0x6A09E667u, c1 = 0xBB67AE85u, c2 = 0x3C6EF372u, c3 = 0xA54FF53Au;
  uint32_t c4 = 0x510E527Fu, c5 = 0x9B05688Cu, c6 = 0x1F83D9ABu, c7 = 0x5BE0CD19u;
  uint32_t c8 = 0x6A09E667u, c9 = 0xBB67AE85u, c10 = 0x3C6EF372u, c11 = 0xA54FF53Au;
  uint32_t c12 = 0, c13 = 0, c14 = 64, c15 = B3_CHUNK_START | B3_CHUNK_END | B3_ROOT;
  FLK_B3_G(c0, c4, c8, c12, dig[0], dig[1]);
  FLK_B3_G(c1, c5, c9, c13, dig[2], dig[3]);
  FLK_B3_G(c2, c6, c10, c14, dig[4], dig[5]);
  FLK_B3_G(c3, c7, c11, c15, dig[6], dig[7]);
  for (uint64_t k = 0; k < sweeps; ++k) {
    const uint64_t base = start + k * T;
    // every nonce of this sweep is >= base: nothing here can beat the minimum
    if (base > __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    if (tid < end - base) {  // (base + tid < end without wrapping past 2^64)
      const uint64_t n = base + tid;
      const uint32_t m[16] = {dig[0], dig[1], dig[2], dig[3], dig[4], dig[5], dig[6], dig[7],
                              (uint32_t)n, (uint32_t)(n >> 32), 0, 0, 0, 0, 0, 0};
      uint32_t v0 = c0, v1 = c1, v2 = c2, v3 = c3, v4 = c4, v5 = c5, v6 = c6, v7 = c7;
      uint32_t v8 = c8, v9 = c9, v10 = c10, v11 = c11, v12 = c12, v13 = c13, v14 = c14, v15 = c15;
      // round 1's diagonal step, then rounds 2..7 (message schedules as in b3_compress_cv)
      FLK_B3_G(v0, v5, v10, v15, m[8], m[9]);
      FLK_B3_G(v1, v6, v11, v12, m[10], m[11]);
      FLK_B3_G(v2, v7, v8, v13, m[12], m[13]);
      FLK_B3_G(v3, v4, v9, v14, m[14], m[15]);
      FLK_B3_ROUND(2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8);
      FLK_B3_ROUND(3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1);
      FLK_B3_ROUND(10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6);
      FLK_B3_ROUND(12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4);
      FLK_B3_ROUND(9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7);
      FLK_B3_ROUND(11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13);
      // output words 0 and 1, read big-endian (byte 0 first)
      const uint64_t be = ((uint64_t)__builtin_bswap32(v0 ^ v8) << 32) | __builtin_bswap32(v1 ^ v9);
      const bool ok = bits == 0 || (be >> (64 - bits)) == 0;
      if (ok) atomicMin(best, (unsigned long long)n);
    }
  }
}

struct PowDigest {
  uint4 lo, hi;
};

PowDigest digest_words(const uint8_t* d) {
  uint32_t w[8];
  for (int i = 0; i < 8; ++i)
    w[i] = (uint32_t)d[4 * i] | ((uint32_t)d[4 * i + 1] << 8) | ((uint32_t)d[4 * i + 2] << 16) |
           ((uint32_t)d[4 * i + 3] << 24);
  return PowDigest{make_uint4(w[0], w[1], w[2], w[3]), make_uint4(w[4], w[5], w[6], w[7])};
}

// Smallest match in [start, start + len) (len capped so the range ends at
// 2^64 - 1), or NONE; `best` is device scratch on `st`.
// `max_grid` work-groups of POW_WG lanes per sweep.
int scan_range(const PowDigest& dg, uint64_t start, uint64_t len, uint32_t bits, unsigned long long* best,
               unsigned long long* host, hipStream_t st, unsigned max_grid, uint64_t* out) {
  const uint64_t end = len > NONE - start ? NONE : start + len;
  hipLaunchKernelGGL(pow_init_kernel, dim3(1), dim3(64), 0, st, best);
  FLK_LAUNCH_CHECK("pow_init_kernel");
  if (end > start) {
    const uint64_t lanes = end - start;
    const uint64_t wgs = (lanes + POW_WG - 1) / POW_WG;
    const unsigned grid = (unsigned)(wgs < max_grid ? wgs : max_grid);
    hipLaunchKernelGGL(pow_scan_kernel, dim3(grid), dim3(POW_WG), 0, st, dg.lo, dg.hi, start, end, bits, best);
    FLK_LAUNCH_CHECK("pow_scan_kernel");
  }
  hipError_t e = read_to_host(host, best, sizeof *host, st);  // the mailbox: no stream synchronisation
  if (e != hipSuccess) return device_failure(e, "pow_scan readback");
  *out = *host;
  return FLK_OK;
}

// Per-device scratch of the (synchronous) PoW calls: the 8-byte device
// minimum and a pinned host word to read it into, made once per device and
// held by one call at a time (the lock is taken for the whole call).
struct PowScratch {
  unsigned long long* dev = nullptr;
  unsigned long long* host = nullptr;
};
std::mutex g_pow_mu;
PowScratch g_pow[64];

int pow_scratch(PowScratch** out) {
  int d = 0;
  hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess) return device_failure(e, "hipGetDevice");
  if (d < 0 || d >= 64) return set_error(FLK_E_SHAPE, "pow: device index %d out of range", d);
  PowScratch& ps = g_pow[d];
  if (!ps.dev) {
    e = hipMalloc(reinterpret_cast<void**>(&ps.dev), sizeof(unsigned long long));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ps.dev = nullptr;
      return set_error(FLK_E_NOMEM, "pow: device scratch allocation failed");
    }
    e = hipHostMalloc(reinterpret_cast<void**>(&ps.host), sizeof(unsigned long long), hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(ps.dev);
      ps.dev = nullptr;
      ps.host = nullptr;
      return set_error(FLK_E_NOMEM, "pow: pinned host scratch allocation failed");
    }
  }
  *out = &ps;
  return FLK_OK;
}

}  // namespace
}  // namespace flk

using namespace flk;

extern "C" {

int flk_pow_scan(const uint8_t* state_digest, uint64_t start, uint64_t len, uint32_t bits, uint64_t* nonce_out,
                 void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!state_digest || !nonce_out) return set_error(FLK_E_SHAPE, "pow_scan: null pointer");
  if (bits > 64) return set_error(FLK_E_SHAPE, "pow_scan: bits=%u > 64", bits);
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_pow_mu);
  PowScratch* ps = nullptr;
  if ((rc = pow_scratch(&ps))) return rc;
  return scan_range(digest_words(state_digest), start, len, bits, ps->dev, ps->host, st, POW_MAX_GRID, nonce_out);
}

int flk_pow_grind(const uint8_t* state_digest, uint32_t bits, uint64_t* nonce_out, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!state_digest || !nonce_out) return set_error(FLK_E_SHAPE, "pow_grind: null pointer");
  if (bits > 48) return set_error(FLK_E_SHAPE, "pow_grind: bits=%u > 48", bits);
  if (bits == 0) {  // the canonical zero-bit nonce (CF/challenger.rs:402-403)
    *nonce_out = 0;
    return FLK_OK;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_pow_mu);
  PowScratch* ps = nullptr;
  if ((rc = pow_scratch(&ps))) return rc;
  const PowDigest dg = digest_words(state_digest);
  // windows of 8x the expected work: a window without a match (p ~ e^-8)
  // moves the search to the next one
  const int wlog = bits + 3 < 14 ? 14 : (bits + 3 > 44 ? 44 : (int)bits + 3);
  const uint64_t window = uint64_t(1) << wlog;
  // sweeps of about half the expected work (at least 2^14 lanes), so the
  // sweep that holds the first match ends soon after it
  const int slog = bits < 15 ? 14 : (bits - 1 > 21 ? 21 : (int)bits - 1);
  const unsigned grid = (unsigned)((uint64_t(1) << slog) / POW_WG);
  uint64_t found = NONE;
  for (uint64_t start = 0; rc == FLK_OK && found == NONE; start += window) {
    rc = scan_range(dg, start, window, bits, ps->dev, ps->host, st, grid, &found);
    if (start > NONE - window) break;  // the whole nonce space is scanned
  }
  if (rc) return rc;
  if (found == NONE) return set_error(FLK_E_SHAPE, "pow_grind: no nonce with %u leading zero bits", bits);
  *nonce_out = found;
  return FLK_OK;
}

}  // extern "C"
