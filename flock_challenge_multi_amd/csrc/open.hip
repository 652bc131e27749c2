// Ligerito opening glue on the device (CF = crates/flock-core/src/):
//
//   flk_gather_rows          the opened rows of the queried positions: row p =
//                            codeword[p * row_len .. +row_len) of the SoA codeword
//                            (CF/ntt/additive_ntt_f128.rs:2700-2720; the query
//                            gathers of CF/pcs/ligerito.rs)
//   flk_merkle_multi_proof   merkle_multi_proof (CF/merkle.rs:1020-1060): the
//                            sibling indices are index arithmetic on the host
//                            (merkle_multi_proof_sibling_indices, :1067-1107),
//                            the hashes a device gather from the tree
//
// Both gathers are random 16-byte reads written out densely; the index lists
// are validated on the host before any launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace flk {
namespace {

constexpr int GATHER_WG = 256;

// dst[t] = src[idx[t / per] * per + t % per], t < n * per (16-byte items).
__global__ __launch_bounds__(GATHER_WG) void gather_kernel(const uint4* __restrict__ src,
                                                           const uint64_t* __restrict__ idx, uint64_t n, uint32_t per,
                                                           uint4* __restrict__ dst) {
  const uint64_t total = n * per;
  for (uint64_t t = (uint64_t)blockIdx.x * GATHER_WG + threadIdx.x; t < total; t += (uint64_t)gridDim.x * GATHER_WG) {
    const uint64_t item = t / per, k = t % per;
    dst[t] = src[idx[item] * per + k];
  }
}

// Copies idx to the device on `st`, gathers, and waits for the stream (the
// host index list must outlive the copy).
int gather(const uint4* src, const std::vector<uint64_t>& idx, uint32_t per, uint4* dst, hipStream_t st,
           const char* what) {
  if (idx.empty()) return FLK_OK;
  uint64_t* d_idx = nullptr;
  hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&d_idx), idx.size() * sizeof(uint64_t), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "%s: index scratch allocation failed", what);
  }
  e = upload_staged(d_idx, idx.data(), idx.size() * sizeof(uint64_t), st);
  if (e != hipSuccess) return device_failure(e, what);
  const uint64_t total = idx.size() * per;
  const uint64_t wgs = (total + GATHER_WG - 1) / GATHER_WG;
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)std::min<uint64_t>(wgs, 4096)), dim3(GATHER_WG), 0, st, src, d_idx,
                     (uint64_t)idx.size(), per, dst);
  FLK_LAUNCH_CHECK("gather_kernel");
  e = flk::stream_free(d_idx, st);
  if (e != hipSuccess) return device_failure(e, what);
  return FLK_OK;
}

}  // namespace
}  // namespace flk

using namespace flk;

extern "C" {

int flk_gather_rows(const flk_f128* codeword, size_t n_rows, size_t row_len, const uint64_t* positions, size_t n,
                    flk_f128* out, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (n == 0) return FLK_OK;
  if (!codeword || !positions || !out) return set_error(FLK_E_SHAPE, "gather_rows: null pointer");
  if (row_len == 0 || row_len > (1u << 20)) return set_error(FLK_E_SHAPE, "gather_rows: row_len=%zu", row_len);
  std::vector<uint64_t> idx(positions, positions + n);
  for (size_t i = 0; i < n; ++i)
    if (idx[i] >= n_rows) return set_error(FLK_E_SHAPE, "gather_rows: position %llu >= %zu rows",
                                           (unsigned long long)idx[i], n_rows);
  return gather(reinterpret_cast<const uint4*>(codeword), idx, (uint32_t)row_len, reinterpret_cast<uint4*>(out),
                static_cast<hipStream_t>(stream), "gather_rows");
}

int flk_merkle_multi_proof(const uint8_t* tree, size_t num_leaves, const uint64_t* positions, size_t n,
                           uint8_t* proof_out, size_t max_hashes, size_t* n_hashes_out, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!n_hashes_out || (n && !positions)) return set_error(FLK_E_SHAPE, "merkle_multi_proof: null pointer");
  if (!is_pow2(num_leaves)) return set_error(FLK_E_SHAPE, "merkle_multi_proof: num_leaves=%zu", num_leaves);
  *n_hashes_out = 0;
  if (n == 0 || num_leaves == 1) return FLK_OK;
  if (!tree || !proof_out) return set_error(FLK_E_SHAPE, "merkle_multi_proof: null pointer");
  std::vector<uint64_t> act(positions, positions + n);
  std::sort(act.begin(), act.end());
  act.erase(std::unique(act.begin(), act.end()), act.end());
  if (act.back() >= num_leaves)
    return set_error(FLK_E_SHAPE, "merkle_multi_proof: position %llu >= %zu leaves", (unsigned long long)act.back(),
                     num_leaves);
  // level by level: a position whose sibling is not active emits the sibling
  std::vector<uint64_t> sib;
  uint64_t level_start = 0;
  for (uint64_t level_len = num_leaves; level_len > 1; level_len >>= 1) {
    size_t w = 0;
    for (size_t i = 0; i < act.size();) {
      const uint64_t p = act[i];
      if (i + 1 < act.size() && act[i + 1] == (p ^ 1)) {
        i += 2;
      } else {
        sib.push_back(level_start + (p ^ 1));
        i += 1;
      }
      act[w++] = p >> 1;
    }
    act.resize(w);
    level_start += level_len;
  }
  if (sib.size() > max_hashes)
    return set_error(FLK_E_SHAPE, "merkle_multi_proof: %zu hashes > capacity %zu", sib.size(), max_hashes);
  rc = gather(reinterpret_cast<const uint4*>(tree), sib, 2, reinterpret_cast<uint4*>(proof_out),
              static_cast<hipStream_t>(stream), "merkle_multi_proof");
  if (rc) return rc;
  *n_hashes_out = sib.size();
  return FLK_OK;
}

}  // extern "C"
