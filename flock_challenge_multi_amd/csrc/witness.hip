// BLAKE3 R1CS witness generation on the device (CP = crates/flock-prover/src/):
//
//   flk_blake3_witness   generate_witness_with_ab_packed_and_lincheck
//                            CP/r1cs_hashes/blake3.rs:4617-4647
//                        (z_lincheck == NULL: generate_witness_with_ab_packed,
//                            :2184-2206)
//   batch layout         drive_witness_packed_and_lincheck, drive_witness_packed
//                            CP/r1cs_hashes/common.rs:196-324,327-360
//
// Output: for each of the 2^n_blocks_log slots, the K = 2^14-bit row vectors
// z, a = A z and b = B z of one compression (2 KiB each, 128 F128), slot j at
// byte offset j * 2048; slots past n_blocks hold the all-zero compression.
// The lincheck stripe (optional) is the 8-slot bit transpose of z: byte
// g * 2^14 + p has bit r = z bit p of slot 8 g + r.
//
// One lane builds one compression's rows (witness_lane.hpp); a work-group is
// one wave of 64 compressions, i.e. 8 stripe groups.  All lanes emit the same
// word positions at the same time, so the words go to a per-lane ring of 20
// u64 per row vector in LDS, and at each 16-word boundary the wave writes the
// window out: per row vector 8 stores of 16 bytes per lane, each store
// covering the whole 128-byte lines of 8 blocks, and the window's stripe bytes
// (two 64-byte chunks per lane) transposed from the same LDS words.  The
// compression runs twice per lane: once for out_lo, whose rows (words 4..7)
// precede the G records, and once while the rows stream out.  The kernel is
// bound by HBM writes: 8 KiB per compression with the stripe, 6 KiB without.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "internal.hpp"
#include "witness_lane.hpp"

namespace flk {
namespace {

constexpr int WIT_WG = 64;              // compressions per work-group (one wave)
constexpr int RING = 20;                // ring slots per lane and row vector
constexpr int RSTRIDE = RING + 1;       // odd lane stride: ring writes hit every bank once per 32 lanes
constexpr int ARR = WIT_WG * RSTRIDE;   // u64 per row-vector ring
constexpr int WIN = 16;                 // words per output window (128 bytes of each block)
constexpr int F128_PER_BLOCK = wit::WORDS / 2;
constexpr int STRIPE_U4_PER_GROUP = wit::K_BITS / 16;

static_assert(WIN + 4 <= RING, "a G call emits at most 4 words after a checkpoint");
static_assert(sizeof(flk_compression) == 112, "compression record is 7 x 16 bytes");

// Message word i of the lane's compression, from LDS ([16][64]: conflict-free).
struct LdsMsg {
  const uint32_t* m;
  int lane;
  __device__ __forceinline__ uint32_t operator()(int i) const { return m[i * WIT_WG + lane]; }
};

struct WaveSink {
  uint64_t* ring;  // LDS [3][64][RSTRIDE]
  int lane;
  int hi;          // end of the part's words
  int flushed;     // first word not yet written out (a multiple of WIN)
  uint64_t blk0, n_total;
  uint4 *z, *a, *b, *stripe;

  // Words outside the part are stored too (no branch): a part's first step
  // writes them before its first word, in slots of no pending word, and its
  // last step after its last word, in slots of words already written out.
  __device__ __forceinline__ void put(int widx, uint64_t zw, uint64_t aw, uint64_t bw) {
    uint64_t* p = ring + lane * RSTRIDE + widx % RING;
    p[0] = zw;
    p[ARR] = aw;
    p[2 * ARR] = bw;
  }
  // The ring is idle before a part's first word: its first 4 KB hold the
  // state words while the part fast-forwards ([16][64] u32, own lane only).
  __device__ __forceinline__ wit::StridedState state_scratch() {
    return wit::StridedState{reinterpret_cast<uint32_t*>(ring) + lane, WIT_WG};
  }
  __device__ __forceinline__ void state_scratch_done() { __syncthreads(); }
  __device__ __forceinline__ bool checkpoint(int words) {
    if (words >= flushed + WIN && flushed < hi) {
      flush(flushed);
      flushed += WIN;
    }
    return flushed >= hi;
  }
  // Write words [w0, w0 + WIN) of all 64 blocks, and their stripe bytes.
  __device__ void flush(int w0) {
    __syncthreads();
    const int q = lane & 7;  // word pair within the window
    const int slot = (w0 + 2 * q) % RING;  // even, so slot + 1 < RING
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      uint4* out = v == 0 ? z : v == 1 ? a : b;
      const uint64_t* src = ring + v * ARR + slot;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int bl = (lane >> 3) + 8 * i;
        const uint64_t blk = blk0 + bl;
        const uint64_t x = src[bl * RSTRIDE], y = src[bl * RSTRIDE + 1];
        if (blk < n_total)
          out[blk * F128_PER_BLOCK + (w0 >> 1) + q] =
              make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
      }
    }
    if (stripe) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int t = lane + WIT_WG * i;
        const int g = t >> 4, w = t & 15;
        const uint64_t grp = blk0 / 8 + g;
        const uint64_t* src = ring + (w0 + w) % RING;
        uint64_t wv[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) wv[r] = src[(8 * g + r) * RSTRIDE];
        uint32_t o[16];
        wit::stripe_chunk(wv, o);
        if (grp * 8 < n_total) {
          uint4* dst = stripe + grp * STRIPE_U4_PER_GROUP + (uint64_t)(w0 + w) * 4;
#pragma unroll
          for (int h = 0; h < 4; ++h) dst[h] = make_uint4(o[4 * h], o[4 * h + 1], o[4 * h + 2], o[4 * h + 3]);
        }
      }
    }
    __syncthreads();
  }
};

// Work-group (c, p) = block index c * parts + p builds words [256 p / parts,
// 256 (p + 1) / parts) of compressions 64 c .. 64 c + 63.
__global__ __launch_bounds__(WIT_WG) void blake3_witness_kernel(const uint4* __restrict__ comps, uint64_t n_blocks,
                                                                uint64_t n_total, int log_parts,
                                                                uint4* __restrict__ z, uint4* __restrict__ a,
                                                                uint4* __restrict__ b, uint4* __restrict__ stripe) {
  __shared__ uint64_t ring[3 * ARR];
  __shared__ uint32_t msg[16 * WIT_WG];
  const int lane = threadIdx.x;
  const int part = blockIdx.x & ((1 << log_parts) - 1);
  const uint64_t blk0 = (uint64_t)(blockIdx.x >> log_parts) * WIT_WG;
  const uint64_t blk = blk0 + lane;

  uint32_t w[28];
#pragma unroll
  for (int i = 0; i < 28; ++i) w[i] = 0;  // padding slots: the all-zero compression
  if (blk < n_blocks) {
    const uint4* src = comps + blk * 7;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const uint4 v = src[i];
      w[4 * i] = v.x;
      w[4 * i + 1] = v.y;
      w[4 * i + 2] = v.z;
      w[4 * i + 3] = v.w;
    }
  }
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cv[i] = w[i];
#pragma unroll
  for (int i = 0; i < 16; ++i) msg[i * WIT_WG + lane] = w[8 + i];
  const uint32_t t_lo = w[24], t_hi = w[25], blen = w[26], flags = w[27];

  const LdsMsg mm{msg, lane};
  const int span = wit::WORDS >> log_parts;
  WaveSink sink{ring, lane, span * (part + 1), span * part, blk0, n_total, z, a, b, stripe};
  wit::block_rows(cv, mm, t_lo, t_hi, blen, flags, span * part, sink);
}

// Parts per compression group: enough work-groups to give every SIMD one
// (4 per CU on 256 CUs) when the batch is small, one part when it is large.
// FLOCK_WIT_PARTS overrides (1, 2, 4, 8 or 16).
int witness_log_parts(size_t groups) {
  static const int env = [] {
    const char* e = getenv("FLOCK_WIT_PARTS");
    const int v = e ? atoi(e) : 0;
    return (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) ? v : 0;
  }();
  int parts = env;
  if (!parts) {
    parts = 1;
    while (parts < 4 && groups * parts < 1024) parts *= 2;
  }
  int l = 0;
  while ((1 << l) < parts) ++l;
  return l;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace
}  // namespace flk

using namespace flk;

extern "C" {

int flk_blake3_witness(const flk_compression* blocks, size_t n_blocks, size_t n_blocks_log, flk_f128* z,
                       flk_f128* a, flk_f128* b, uint8_t* z_lincheck, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (n_blocks_log < 3 || n_blocks_log > 30)
    return set_error(FLK_E_SHAPE, "blake3_witness: n_blocks_log=%zu must be in 3..30", n_blocks_log);
  const size_t n_total = size_t(1) << n_blocks_log;
  if (n_blocks > n_total)
    return set_error(FLK_E_SHAPE, "blake3_witness: %zu compressions > 2^%zu slots", n_blocks, n_blocks_log);
  if ((n_blocks && !blocks) || !z || !a || !b) return set_error(FLK_E_SHAPE, "blake3_witness: null pointer");
  if ((n_blocks && !aligned16(blocks)) || !aligned16(z) || !aligned16(a) || !aligned16(b) ||
      (z_lincheck && !aligned16(z_lincheck)))
    return set_error(FLK_E_SHAPE, "blake3_witness: buffers must be 16-byte aligned");
  const size_t groups = (n_total + WIT_WG - 1) / WIT_WG;
  const int log_parts = witness_log_parts(groups);
  const unsigned grid = (unsigned)(groups << log_parts);
  hipLaunchKernelGGL(blake3_witness_kernel, dim3(grid), dim3(WIT_WG), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const uint4*>(blocks), (uint64_t)n_blocks, (uint64_t)n_total, log_parts,
                     reinterpret_cast<uint4*>(z), reinterpret_cast<uint4*>(a), reinterpret_cast<uint4*>(b),
                     reinterpret_cast<uint4*>(z_lincheck));
  FLK_LAUNCH_CHECK("blake3_witness_kernel");
  return FLK_OK;
}

}  // extern "C"
