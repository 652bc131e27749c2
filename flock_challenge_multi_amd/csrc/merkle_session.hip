// Streaming Merkle session: the C-ABI counterpart of crate::gpu::merkle's
// begin / Session::{commit_leaves, commit_parent_levels, finish}
// (CF/gpu.rs:1228-1490) and of crate::gpu::prewire (CF/gpu.rs:497-515).
//
// The Metal session binds the caller's leaf bytes and tree memory zero-copy.
// Discrete HBM cannot alias host memory, so here each buffer is classified
// once at begin:
//   * device memory (hipMalloc / flk_alloc) is used in place;
//   * host memory is mirrored in a session-owned device buffer: leaf ranges are
//     copied up as they are committed, and finish copies every node the device
//     wrote back into the caller's tree.
// The hybrid CPU join of the reference (the caller writes some leaf CVs into the
// tree region itself before commit_parent_levels, CF/gpu.rs:1161-1170) is kept:
// with a host tree, commit_parent_levels first uploads the leaf slots no
// commit_leaves call covered.  All work of a session runs in order on the
// stream given to begin (or the null stream).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "internal.hpp"

namespace flk {
int merkle_leaves_device(const uint8_t* data, size_t leaf_size, size_t n_leaves, uint8_t* out, hipStream_t st);
int merkle_upper_levels_device(uint8_t* tree_bytes, size_t n, size_t stop_nodes, hipStream_t st);
}  // namespace flk

struct flk_merkle_session {
  const uint8_t* data_host = nullptr;  // caller's bytes when they are host memory
  const uint8_t* data_dev = nullptr;   // bytes the kernels read
  uint8_t* data_own = nullptr;         // session-owned mirror of host bytes
  uint8_t* tree_host = nullptr;        // caller's tree when it is host memory
  uint8_t* tree_dev = nullptr;         // tree the kernels write
  uint8_t* tree_own = nullptr;         // session-owned mirror of a host tree
  size_t n_leaves = 0, leaf_size = 0, stop_nodes = 0, needed_nodes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  bool started = false, failed = false, parents_done = false;
  std::vector<std::pair<size_t, size_t>> gpu_leaves;  // committed [lo, hi) ranges
};

namespace {

using flk::set_error;

bool is_device_pointer(const void* p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory reports an error on some runtimes
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

// Leaves plus every parent level with count >= stop (CF/gpu.rs needed_nodes).
size_t needed_nodes(size_t n_leaves, size_t stop) {
  size_t total = n_leaves;
  for (size_t count = n_leaves / 2; count >= stop && count >= 1; count /= 2) total += count;
  return total;
}

void release(flk_merkle_session* s) {
  if (!s) return;
  if (s->data_own) (void)hipFree(s->data_own);
  if (s->tree_own) (void)hipFree(s->tree_own);
  if (s->ev_start) (void)hipEventDestroy(s->ev_start);
  if (s->ev_end) (void)hipEventDestroy(s->ev_end);
  delete s;
}

int fail(flk_merkle_session* s, hipError_t e, const char* what) {
  s->failed = true;
  return flk::device_failure(e, what);
}

int mark_start(flk_merkle_session* s) {
  if (s->started) return FLK_OK;
  const hipError_t e = hipEventRecord(s->ev_start, s->stream);
  if (e != hipSuccess) return fail(s, e, "merkle session: event record");
  s->started = true;
  return FLK_OK;
}

}  // namespace

extern "C" {

int flk_prewire(const void* data, size_t len) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!data && len) return set_error(FLK_E_SHAPE, "prewire: null pointer");
  // The Metal facade pre-wraps large host buffers for zero-copy use; host
  // memory registration is not used by this library (host buffers are staged
  // by HIP's copy engine), so there is nothing to prepare.
  return FLK_OK;
}

int flk_merkle_begin(const uint8_t* data, size_t data_len, size_t leaf_size, uint8_t* tree, size_t tree_len,
                     size_t stop_nodes, void* stream, flk_merkle_session** out) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!out) return set_error(FLK_E_SHAPE, "merkle_begin: null session out-pointer");
  *out = nullptr;
  if (!data || !tree) return set_error(FLK_E_SHAPE, "merkle_begin: null pointer");
  if (leaf_size == 0 || leaf_size % 16 || leaf_size > 1024)
    return set_error(FLK_E_SHAPE, "merkle_begin: leaf size %zu (multiples of 16 up to 1024)", leaf_size);
  if (data_len == 0 || data_len % leaf_size)
    return set_error(FLK_E_SHAPE, "merkle_begin: %zu bytes is not a whole number of %zu-byte leaves", data_len,
                     leaf_size);
  if ((uintptr_t)data % 16 || (uintptr_t)tree % 16)
    return set_error(FLK_E_SHAPE, "merkle_begin: data and tree must be 16-byte aligned");
  const size_t n = data_len / leaf_size;
  const size_t stop = stop_nodes < 1 ? 1 : stop_nodes;
  const bool parents = stop <= n / 2;
  if (parents && !flk::is_pow2(n))
    return set_error(FLK_E_SHAPE, "merkle_begin: %zu leaves; parent levels need a power of two", n);
  const size_t needed = parents ? needed_nodes(n, stop) : n;
  if (tree_len < needed)
    return set_error(FLK_E_SHAPE, "merkle_begin: tree holds %zu nodes, the session writes %zu", tree_len, needed);

  flk_merkle_session* s = new (std::nothrow) flk_merkle_session;
  if (!s) return set_error(FLK_E_NOMEM, "merkle_begin: out of host memory");
  s->n_leaves = n;
  s->leaf_size = leaf_size;
  s->stop_nodes = stop;
  s->needed_nodes = needed;
  s->stream = static_cast<hipStream_t>(stream);
  hipError_t e = hipEventCreate(&s->ev_start);
  if (e == hipSuccess) e = hipEventCreate(&s->ev_end);
  if (e != hipSuccess) {
    release(s);
    return flk::device_failure(e, "merkle_begin: event create");
  }
  if (is_device_pointer(data)) {
    s->data_dev = data;
  } else {
    s->data_host = data;
    if (hipMalloc(reinterpret_cast<void**>(&s->data_own), data_len) != hipSuccess) {
      (void)hipGetLastError();
      release(s);
      return set_error(FLK_E_NOMEM, "merkle_begin: no device memory for %zu leaf bytes", data_len);
    }
    s->data_dev = s->data_own;
  }
  if (is_device_pointer(tree)) {
    s->tree_dev = tree;
  } else {
    s->tree_host = tree;
    if (hipMalloc(reinterpret_cast<void**>(&s->tree_own), needed * 32) != hipSuccess) {
      (void)hipGetLastError();
      release(s);
      return set_error(FLK_E_NOMEM, "merkle_begin: no device memory for %zu tree nodes", needed);
    }
    s->tree_dev = s->tree_own;
  }
  *out = s;
  return FLK_OK;
}

size_t flk_merkle_session_leaves(const flk_merkle_session* s) { return s ? s->n_leaves : 0; }

int flk_merkle_commit_leaves(flk_merkle_session* s, size_t leaf_lo, size_t leaf_hi) {
  if (!s) return set_error(FLK_E_SHAPE, "commit_leaves: null session");
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (s->failed) return set_error(FLK_E_DISABLED, "commit_leaves: the session has failed");
  if (s->parents_done) return set_error(FLK_E_SHAPE, "commit_leaves: parent levels were already committed");
  if (leaf_lo >= leaf_hi || leaf_hi > s->n_leaves)
    return set_error(FLK_E_SHAPE, "commit_leaves: range [%zu, %zu) of %zu leaves", leaf_lo, leaf_hi, s->n_leaves);
  rc = mark_start(s);
  if (rc) return rc;
  const size_t off = leaf_lo * s->leaf_size, bytes = (leaf_hi - leaf_lo) * s->leaf_size;
  if (s->data_host) {
    const hipError_t e = hipMemcpyAsync(s->data_own + off, s->data_host + off, bytes, hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return fail(s, e, "commit_leaves: leaf upload");
  }
  rc = flk::merkle_leaves_device(s->data_dev + off, s->leaf_size, leaf_hi - leaf_lo, s->tree_dev + 32 * leaf_lo,
                                 s->stream);
  if (rc) {
    s->failed = true;
    return rc;
  }
  s->gpu_leaves.emplace_back(leaf_lo, leaf_hi);
  return FLK_OK;
}

int flk_merkle_commit_parent_levels(flk_merkle_session* s) {
  if (!s) return set_error(FLK_E_SHAPE, "commit_parent_levels: null session");
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (s->failed) return set_error(FLK_E_DISABLED, "commit_parent_levels: the session has failed");
  if (s->parents_done) return set_error(FLK_E_SHAPE, "commit_parent_levels: called twice");
  s->parents_done = true;
  if (s->stop_nodes > s->n_leaves / 2) return FLK_OK;  // leaf-only session
  rc = mark_start(s);
  if (rc) return rc;
  if (s->tree_host) {
    // Leaf slots no commit_leaves call covered hold the caller's CPU-join CVs.
    std::vector<std::pair<size_t, size_t>> r = s->gpu_leaves;
    std::sort(r.begin(), r.end());
    size_t at = 0;
    r.emplace_back(s->n_leaves, s->n_leaves);
    for (const auto& [lo, hi] : r) {
      if (lo > at) {
        const hipError_t e = hipMemcpyAsync(s->tree_own + 32 * at, s->tree_host + 32 * at, 32 * (lo - at),
                                            hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) return fail(s, e, "commit_parent_levels: CPU-join leaf upload");
      }
      at = std::max(at, hi);
    }
  }
  rc = flk::merkle_upper_levels_device(s->tree_dev, s->n_leaves, s->stop_nodes, s->stream);
  if (rc) s->failed = true;
  return rc;
}

int flk_merkle_finish(flk_merkle_session* s, double* gpu_seconds) {
  if (!s) return set_error(FLK_E_SHAPE, "merkle_finish: null session");
  int rc = s->failed ? set_error(FLK_E_DISABLED, "merkle_finish: the session has failed") : FLK_OK;
  hipError_t e = hipSuccess;
  if (!rc && s->tree_host) {
    for (const auto& [lo, hi] : s->gpu_leaves) {
      e = hipMemcpyAsync(s->tree_host + 32 * lo, s->tree_own + 32 * lo, 32 * (hi - lo), hipMemcpyDeviceToHost,
                         s->stream);
      if (e != hipSuccess) break;
    }
    if (e == hipSuccess && s->parents_done && s->needed_nodes > s->n_leaves)
      e = hipMemcpyAsync(s->tree_host + 32 * s->n_leaves, s->tree_own + 32 * s->n_leaves,
                         32 * (s->needed_nodes - s->n_leaves), hipMemcpyDeviceToHost, s->stream);
  }
  if (!rc && e == hipSuccess && s->started) e = hipEventRecord(s->ev_end, s->stream);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(s->stream);
  float ms = 0.f;
  if (!rc && e == hipSuccess && s->started) e = hipEventElapsedTime(&ms, s->ev_start, s->ev_end);
  if (!rc && e != hipSuccess) rc = flk::device_failure(e, "merkle_finish");
  if (!rc && gpu_seconds) *gpu_seconds = ms * 1e-3;
  if (rc) (void)hipStreamSynchronize(s->stream);  // nothing may still write session memory
  release(s);
  return rc;
}

void flk_merkle_discard(flk_merkle_session* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  release(s);
}

}  // extern "C"
