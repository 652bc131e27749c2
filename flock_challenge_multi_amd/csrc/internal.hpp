// Shared host-side plumbing of the device library: status codes, the
// process-wide failure latch, the per-device context (cached twiddle tables)
// and launch checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/flock_gpu.h"
#include "host_api.hpp"

namespace flk {

// Record a message for flk_last_error() and return `code`.
int set_error(int code, const char* fmt, ...);
// A runtime failure: latch the library disabled (CF/gpu.rs:455-466) and record.
int device_failure(hipError_t e, const char* where);
// FLK_OK unless the library is disabled (latch, API or FLOCK_NO_GPU).
int check_enabled();
// Device copy of the twiddle table of AdditiveNttF128::standard(dim), cached per
// (device, dim).  Returns nullptr (and records the error) on failure.
const uint4* twiddles_device(size_t dim, int* status);
// Host copy of the same table (computed once per dim).
const flk_f128* twiddles_host(size_t dim);

inline bool is_pow2(size_t n) { return n && !(n & (n - 1)); }
inline int log2_exact(size_t n) {
  if (!is_pow2(n)) return -1;
  int l = 0;
  while ((size_t(1) << l) < n) ++l;
  return l;
}

#define FLK_LAUNCH_CHECK(where)                              \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) return ::flk::device_failure(e_, where); \
  } while (0)

// Stream-ordered device scratch.  Inside an ArenaScope the calling thread's
// allocations come from that arena (one device block carved by a best-fit free
// list, set up once and reused: no HIP call per allocation); elsewhere they come
// from the stream-ordered pool (hipMallocAsync / hipFreeAsync).  Arena blocks are
// handed out again as soon as they are freed, which is safe because everything
// in an arena scope is ordered on one stream.  stream_free recognises arena
// blocks of any live arena, in or out of the scope.
hipError_t stream_alloc(void** p, size_t bytes, hipStream_t st);
// Copy bytes from the device to host memory and wait for st (a pinned bounce
// buffer per thread for reads up to 1 MiB).
hipError_t read_to_host(void* dst, const void* src, size_t bytes, hipStream_t st);
// Host -> device copy enqueued on st through the calling thread's pinned upload
// ring: src may be reused as soon as this returns, and nothing waits for the
// copy (uploads above 8 MiB are copied and waited for).
hipError_t upload_staged(void* dst, const void* src, size_t bytes, hipStream_t st);
// p[0 .. n) = 0 by a kernel on st (graph-capturable, unlike hipMemsetAsync).
int zero_n(uint4* p, uint64_t n, hipStream_t st);
hipError_t stream_free(void* p, hipStream_t st);

// Work-group partial sums of the reduction kernels.  A kernel launched on a grid
// of G work-groups stores the XOR over its work-group of output k at
// parts[k * G + blockIdx.x] (plain stores); partials_sum_free then writes
// out[k] = XOR of the G partials and releases the scratch.  The scratch comes
// from the stream-ordered pool (hipMallocAsync / hipFreeAsync on the call's
// stream), so calls on different streams never share it.  Device-scope atomics
// from every work-group into the same few words cost ~7 ns each, serialised.
// With grid == 1 the partials are out itself (no scratch, no second launch).
int partials_alloc(uint4** parts, unsigned grid, int n_out, uint4* out, hipStream_t st);
int partials_sum_free(uint4* parts, unsigned grid, int n_out, uint4* out, hipStream_t st);
// out[k] = XOR of parts[k * grid .. k * grid + grid) for k < n_out, without freeing
// parts (caller-owned scratch).
int partials_sum(const uint4* parts, unsigned grid, int n_out, uint4* out, hipStream_t st);

// Reductions whose result the calling thread reads back at once.  reduce_begin
// takes the partials (partials_alloc) and, when the calling thread armed a
// read-back of the n_out results (flk_read_arm), its mailbox; reduce_end sums
// the partials on one launch that also stores the results into the mailbox and
// sets its sequence word, so no publish kernel follows (without an arm it is
// partials_sum_free).
struct ReduceTail {
  uint4* parts;   // partials, parts[k * grid + blockIdx.x]
  uint4* out;     // the n_out results (device)
  uint4* box;     // mapped host memory for the results; NULL: no arm taken
  uint32_t* flag; // the mailbox's sequence word
  uint32_t seq;
};
int reduce_begin(unsigned grid, int n_out, uint4* out, hipStream_t st, ReduceTail* rt);
int reduce_end(const ReduceTail& rt, unsigned grid, int n_out, hipStream_t st);

// fold_1b_rows against the split eq tensor eq_lo (2^lo_log) x eq_hi (2^hi_log),
// lo_log >= 8: out128 (device) = sum_i bit_r(z_i) eq_lo[i mod 2^lo_log] eq_hi[i >> lo_log].
int rs_fold_1b_rows_split(const flk_f128* z, const flk_f128* eq_lo, size_t lo_log, const flk_f128* eq_hi,
                          size_t hi_log, flk_f128* out128, hipStream_t st);

// Host-side GF(2^128) product (used for set-up constants only).
flk_f128 host_mul(flk_f128 a, flk_f128 b);
flk_f128 host_inv(flk_f128 a);
flk_f128 phi8_host(uint8_t v);
int lagrange_weights_nodes(size_t k_skip, flk_f128 z, size_t base, flk_f128* out);
void zc_fixed_challenges(flk_f128 out[7]);

}  // namespace flk
