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
// digest words are kernel arguments (scalar registers); the message words
// 10..15 are zero and fold away.  This is compute work: ~56 G functions per
// nonce, no memory traffic beyond one 8-byte minimum.
#include <hip/hip_runtime.h>

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
  uint32_t m[16] = {d_lo.x, d_lo.y, d_lo.z, d_lo.w, d_hi.x, d_hi.y, d_hi.z, d_hi.w, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t k = 0; k < sweeps; ++k) {
    const uint64_t base = start + k * T;
    // every nonce of this sweep is >= base: nothing here can beat the minimum
    if (base > __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    if (tid < end - base) {  // (base + tid < end without wrapping past 2^64)
      const uint64_t n = base + tid;
      m[8] = (uint32_t)n;
      m[9] = (uint32_t)(n >> 32);
      uint32_t cv[8];
      b3_iv(cv);
      b3_compress_cv(cv, m, 0, 0, 64, B3_CHUNK_START | B3_CHUNK_END | B3_ROOT);
      const uint64_t be = ((uint64_t)__builtin_bswap32(cv[0]) << 32) | __builtin_bswap32(cv[1]);
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
int scan_range(const PowDigest& dg, uint64_t start, uint64_t len, uint32_t bits, unsigned long long* best,
               hipStream_t st, uint64_t* out) {
  const uint64_t end = len > NONE - start ? NONE : start + len;
  hipLaunchKernelGGL(pow_init_kernel, dim3(1), dim3(64), 0, st, best);
  FLK_LAUNCH_CHECK("pow_init_kernel");
  if (end > start) {
    const uint64_t lanes = end - start;
    const uint64_t wgs = (lanes + POW_WG - 1) / POW_WG;
    const unsigned grid = (unsigned)(wgs < POW_MAX_GRID ? wgs : POW_MAX_GRID);
    hipLaunchKernelGGL(pow_scan_kernel, dim3(grid), dim3(POW_WG), 0, st, dg.lo, dg.hi, start, end, bits, best);
    FLK_LAUNCH_CHECK("pow_scan_kernel");
  }
  unsigned long long h = NONE;
  hipError_t e = hipMemcpyAsync(&h, best, sizeof h, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return device_failure(e, "pow_scan readback");
  *out = h;
  return FLK_OK;
}

int with_scratch(hipStream_t st, unsigned long long** best) {
  const hipError_t e = hipMallocAsync(reinterpret_cast<void**>(best), sizeof(unsigned long long), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "pow: scratch allocation failed");
  }
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
  unsigned long long* best = nullptr;
  if ((rc = with_scratch(st, &best))) return rc;
  rc = scan_range(digest_words(state_digest), start, len, bits, best, st, nonce_out);
  (void)hipFreeAsync(best, st);
  return rc;
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
  unsigned long long* best = nullptr;
  if ((rc = with_scratch(st, &best))) return rc;
  const PowDigest dg = digest_words(state_digest);
  // windows of 8x the expected work: a window without a match (p ~ e^-8)
  // moves the search to the next one
  const int wlog = bits + 3 < 14 ? 14 : (bits + 3 > 44 ? 44 : (int)bits + 3);
  const uint64_t window = uint64_t(1) << wlog;
  uint64_t found = NONE;
  for (uint64_t start = 0; rc == FLK_OK && found == NONE; start += window) {
    rc = scan_range(dg, start, window, bits, best, st, &found);
    if (start > NONE - window) break;  // the whole nonce space is scanned
  }
  (void)hipFreeAsync(best, st);
  if (rc) return rc;
  if (found == NONE) return set_error(FLK_E_SHAPE, "pow_grind: no nonce with %u leading zero bits", bits);
  *nonce_out = found;
  return FLK_OK;
}

}  // extern "C"
