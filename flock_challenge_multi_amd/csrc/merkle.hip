// BLAKE3 Merkle tree for gfx950 (CF/merkle.rs:15-22,192-203,498,683,753-815)
// and the PCS commit that chains the RS encode into it
// (CF/pcs/commit.rs:187-227, CF/pcs/ligerito.rs:4750-4868).
//
// Leaves: one lane hashes one leaf (a codeword row of num_ntts * 16 bytes) as
// one BLAKE3 chunk: blocks of 64 bytes, CHUNK_START on the first, CHUNK_END on
// the last, chunk counter 0, never ROOT (finalize_non_root).
// Parents: a work-group hashes a 2^LV-leaf subtree level by level in LDS and
// writes every level it produces into the flat tree, so a 2^20-leaf tree
// takes three launches after the leaves.
#include <hip/hip_runtime.h>

#include "blake3_dev.hpp"
#include "internal.hpp"

namespace flk {

int ntt_rs_encode_device(size_t dim, const uint4* msg, size_t n_msg, uint4* cw, size_t n_cw, size_t num_ntts,
                         hipStream_t st);

// 256 lanes (four waves) per work-group: with 2^17 one-KiB leaves the kernel is
// a single round of two waves per SIMD, and 128-lane work-groups ran 20-30%
// longer right after some kernels (the waves did not all start together).
constexpr int LEAF_WG = 256;

__global__ void __launch_bounds__(LEAF_WG) merkle_leaves_kernel(const uint4* __restrict__ data, uint32_t leaf_words4,
                                                               uint32_t leaf_bytes, uint64_t n_leaves,
                                                               uint4* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * LEAF_WG + threadIdx.x;
  if (i >= n_leaves) return;
  const uint4* p = data + i * leaf_words4;
  uint32_t cv[8];
  b3_iv(cv);
  const uint32_t nblocks = (leaf_bytes + 63u) / 64u;
  // Block b + 1 is loaded while block b is compressed (the chain is sequential).
  uint4 nxt[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) nxt[q] = (uint32_t)q < leaf_words4 ? p[q] : make_uint4(0u, 0u, 0u, 0u);
  for (uint32_t b = 0; b < nblocks; ++b) {
    uint32_t m[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      m[4 * q + 0] = nxt[q].x;
      m[4 * q + 1] = nxt[q].y;
      m[4 * q + 2] = nxt[q].z;
      m[4 * q + 3] = nxt[q].w;
    }
    const uint32_t base4 = (b + 1u) * 4u;
#pragma unroll
    for (int q = 0; q < 4; ++q) nxt[q] = base4 + q < leaf_words4 ? p[base4 + q] : make_uint4(0u, 0u, 0u, 0u);
    const uint32_t rem = leaf_bytes - b * 64u;
    const uint32_t blen = rem < 64u ? rem : 64u;
    const uint32_t flags = (b == 0 ? B3_CHUNK_START : 0u) | (b + 1 == nblocks ? B3_CHUNK_END : 0u);
    b3_compress_cv(cv, m, 0u, 0u, blen, flags);
  }
  out[2 * i] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
  out[2 * i + 1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// One work-group turns 2^LV children (starting at child index blockIdx*2^LV of the
// level at `child`) into the LV levels above them.  level_off[k] is the tree
// offset (in hashes) of level k above the child level, k = 1..LV.  Each node is
// hashed by a lane quad (b3_parent_quad): the levels of a tree top are a chain
// of dependent compressions, so four lanes per compression cut the latency.
constexpr int PAR_MAX_LV = 9;  // 512 children, 256 nodes x 4 lanes

struct LevelOffsets {
  uint64_t off[PAR_MAX_LV + 1];
};

__global__ void __launch_bounds__(1024) merkle_levels_kernel(uint4* __restrict__ tree, uint64_t child_off, uint32_t lv,
                                                            LevelOffsets lo) {
  __shared__ uint32_t cvs[2][256 * 8];
  const uint32_t t = threadIdx.x, node = t >> 2, q = t & 3u;
  const uint64_t first_child = (uint64_t)blockIdx.x << lv;
  for (uint32_t k = 1; k <= lv; ++k) {
    const uint32_t n = 1u << (lv - k);  // nodes produced at this level
    if (node < n) {
      uint32_t m[16];
      if (k == 1) {
        const uint4* c = tree + 2 * (child_off + first_child + 2 * node);
        const uint4 a0 = c[0], a1 = c[1], b0 = c[2], b1 = c[3];
        m[0] = a0.x; m[1] = a0.y; m[2] = a0.z; m[3] = a0.w; m[4] = a1.x; m[5] = a1.y; m[6] = a1.z; m[7] = a1.w;
        m[8] = b0.x; m[9] = b0.y; m[10] = b0.z; m[11] = b0.w; m[12] = b1.x; m[13] = b1.y; m[14] = b1.z; m[15] = b1.w;
      } else {
        const uint32_t* src = cvs[k & 1u] + node * 16;  // children 2 node, 2 node + 1, 8 words each
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = src[i];
      }
      const uint2 o = b3_parent_quad(m, q);
      const uint64_t hash = lo.off[k] + ((uint64_t)blockIdx.x << (lv - k)) + node;
      uint32_t* dst = reinterpret_cast<uint32_t*>(tree + 2 * hash);
      dst[q] = o.x;
      dst[4 + q] = o.y;
      uint32_t* next = cvs[(k + 1) & 1u] + node * 8;
      next[q] = o.x;
      next[4 + q] = o.y;
    }
    __syncthreads();
  }
}

__global__ void merkle_pairs_kernel(const uint4* __restrict__ children, uint64_t n_parents, uint4* __restrict__ parents) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_parents) return;
  const uint4 a0 = children[4 * i], a1 = children[4 * i + 1], b0 = children[4 * i + 2], b1 = children[4 * i + 3];
  uint32_t l[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  uint32_t r[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  uint32_t o[8];
  b3_parent(l, r, o);
  parents[2 * i] = make_uint4(o[0], o[1], o[2], o[3]);
  parents[2 * i + 1] = make_uint4(o[4], o[5], o[6], o[7]);
}

// One BLAKE3 chunk per leaf, read as 16-byte words: 16..1024 bytes in steps of 16
// (the last 64-byte block is partial when the size is not a multiple of 64).
static bool leaf_size_ok(size_t leaf_size) { return leaf_size >= 16 && leaf_size <= 1024 && leaf_size % 16 == 0; }

int merkle_leaves_device(const uint8_t* data, size_t leaf_size, size_t n_leaves, uint8_t* out, hipStream_t st) {
  if (!leaf_size_ok(leaf_size))
    return set_error(FLK_E_SHAPE, "merkle: leaf size %zu unsupported (multiples of 16 up to 1024)", leaf_size);
  if ((uintptr_t)data % 16 || (uintptr_t)out % 16) return set_error(FLK_E_SHAPE, "merkle: buffers must be 16-byte aligned");
  if (n_leaves == 0) return FLK_OK;
  const uint64_t blocks = (n_leaves + LEAF_WG - 1) / LEAF_WG;
  hipLaunchKernelGGL(merkle_leaves_kernel, dim3((unsigned)blocks), dim3(LEAF_WG), 0, st,
                     reinterpret_cast<const uint4*>(data), (uint32_t)(leaf_size / 16), (uint32_t)leaf_size,
                     (uint64_t)n_leaves, reinterpret_cast<uint4*>(out));
  FLK_LAUNCH_CHECK("merkle_leaves_kernel");
  return FLK_OK;
}

// Parent levels of a flat tree whose leaves [0, n) are already written, for as
// long as a level has at least stop_nodes nodes (stop_nodes = 1: up to the root;
// the streamed commit leaves the top of the tree to its caller, CF/gpu.rs:1369).
int merkle_upper_levels_device(uint8_t* tree_bytes, size_t n, size_t stop_nodes, hipStream_t st) {
  uint4* tree = reinterpret_cast<uint4*>(tree_bytes);
  const uint64_t stop = stop_nodes < 1 ? 1 : stop_nodes;
  uint64_t child_off = 0, cur = n;
  while (cur > 1 && cur / 2 >= stop) {
    int lv_total = log2_exact(cur);
    while (lv_total > 1 && (cur >> lv_total) < stop) --lv_total;  // last level produced keeps >= stop nodes
    const uint32_t lv = (uint32_t)(lv_total < PAR_MAX_LV ? lv_total : PAR_MAX_LV);
    LevelOffsets lo{};
    uint64_t off = child_off + cur, len = cur;
    for (uint32_t k = 1; k <= lv; ++k) {
      lo.off[k] = off;
      len >>= 1;
      off += len;
    }
    const uint32_t threads = 4u << (lv - 1);  // a lane quad per node of the first level
    const uint64_t blocks = cur >> lv;
    hipLaunchKernelGGL(merkle_levels_kernel, dim3((unsigned)blocks), dim3(threads < 64 ? 64 : threads), 0, st, tree,
                       child_off, lv, lo);
    FLK_LAUNCH_CHECK("merkle_levels_kernel");
    child_off = lo.off[lv];
    cur >>= lv;
  }
  return FLK_OK;
}

int merkle_tree_device(const uint8_t* data, size_t data_len, size_t num_leaves, int hash_kind, uint8_t* tree,
                       hipStream_t st) {
  int rc = check_enabled();
  if (rc) return rc;
  if (hash_kind != FLK_HASH_BLAKE3)
    return set_error(FLK_E_SHAPE, "merkle: only BLAKE3 trees run on the device (hash_kind=%d)", hash_kind);
  if (!data || !tree) return set_error(FLK_E_SHAPE, "merkle: null pointer");
  if (!is_pow2(num_leaves) || data_len % num_leaves)
    return set_error(FLK_E_SHAPE, "merkle: num_leaves %zu must be a power of two dividing %zu", num_leaves, data_len);
  rc = merkle_leaves_device(data, data_len / num_leaves, num_leaves, tree, st);
  if (rc) return rc;
  return merkle_upper_levels_device(tree, num_leaves, 1, st);
}

}  // namespace flk

extern "C" {

int flk_merkle_hash_leaves(const uint8_t* data, size_t leaf_size, size_t n_leaves, uint8_t* out, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!data || !out) return flk::set_error(FLK_E_SHAPE, "hash_leaves: null pointer");
  return flk::merkle_leaves_device(data, leaf_size, n_leaves, out, static_cast<hipStream_t>(stream));
}

int flk_merkle_hash_pairs_level(const uint8_t* children, size_t n_children, uint8_t* parents, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!children || !parents) return flk::set_error(FLK_E_SHAPE, "hash_pairs_level: null pointer");
  if (n_children % 2) return flk::set_error(FLK_E_SHAPE, "hash_pairs_level: odd child count %zu", n_children);
  const uint64_t n = n_children / 2;
  if (n == 0) return FLK_OK;
  hipLaunchKernelGGL(flk::merkle_pairs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const uint4*>(children), n,
                     reinterpret_cast<uint4*>(parents));
  FLK_LAUNCH_CHECK("merkle_pairs_kernel");
  return FLK_OK;
}

int flk_merkle_tree(const uint8_t* data, size_t data_len, size_t num_leaves, int hash_kind, uint8_t* tree_out,
                    void* stream) {
  return flk::merkle_tree_device(data, data_len, num_leaves, hash_kind, tree_out, static_cast<hipStream_t>(stream));
}

int flk_commit(const flk_f128* poly, size_t log_msg_cols, size_t log_num_interleaved, size_t log_inv_rate,
               int hash_kind, flk_f128* codeword_out, uint8_t* tree_out, uint8_t* root_out, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (log_inv_rate < 1) return flk::set_error(FLK_E_SHAPE, "commit: log_inv_rate must be >= 1");
  if (log_msg_cols + log_inv_rate > 24 || log_num_interleaved > 6)
    return flk::set_error(FLK_E_SHAPE, "commit: shape out of range (cols 2^%zu rate 2^-%zu lanes 2^%zu)", log_msg_cols,
                          log_inv_rate, log_num_interleaved);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t lanes = size_t(1) << log_num_interleaved;
  const size_t n_msg = lanes << log_msg_cols;
  const size_t n_pos = size_t(1) << (log_msg_cols + log_inv_rate);
  const size_t n_cw = n_pos * lanes;
  const size_t dim = log_msg_cols + log_inv_rate;
  if (hash_kind != FLK_HASH_BLAKE3)
    return flk::set_error(FLK_E_SHAPE, "commit: only BLAKE3 trees run on the device (hash_kind=%d)", hash_kind);
  rc = flk::ntt_rs_encode_device(dim, reinterpret_cast<const uint4*>(poly), n_msg, reinterpret_cast<uint4*>(codeword_out),
                                 n_cw, lanes, st);
  if (rc) return rc;
  rc = flk::merkle_tree_device(reinterpret_cast<const uint8_t*>(codeword_out), n_cw * 16, n_pos, hash_kind, tree_out, st);
  if (rc) return rc;
  if (root_out) {
    hipError_t e = flk::read_to_host(root_out, tree_out + 32 * (2 * n_pos - 2), 32, st);
    if (e != hipSuccess) return flk::device_failure(e, "commit: root copy");
  }
  return FLK_OK;
}

}  // extern "C"
