/*
 * Oracle (test infrastructure): BLAKE3 restated from the BLAKE3 specification
 * (the reference uses crate `blake3` 1.8.6, Cargo.lock:27-30, not vendored),
 * and the reference's BLAKE3 Merkle tree.
 *
 *   leaf   = non-root chaining value of the leaf bytes
 *            (CF/merkle.rs:192-196, Hasher::update + finalize_non_root)
 *   parent = PARENT-flagged compression of left||right under the IV, non-root
 *            (CF/merkle.rs:198-203, merge_subtrees_non_root)
 *   layout = flat 2n-1 array, leaves first, root last (CF/merkle.rs:15-22,
 *            merkle_tree_sequential :922-950)
 *   proof  = sibling path leaf->root (CF/merkle.rs:960-1010)
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static const uint32_t IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                               0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
static const int PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };
#define CHUNK_LEN 1024

static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void g(uint32_t *s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] = s[a] + s[b] + x;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

void orc_blake3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                         uint32_t block_len, uint32_t flags, uint32_t out16[16]) {
    uint32_t m[16], t[16], s[16];
    for (int i = 0; i < 16; i++)
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) |
               ((uint32_t)block[4 * i + 2] << 16) | ((uint32_t)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = IV[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);
        g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);
        g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);
        g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]);
        g(s, 3, 4, 9, 14, m[14], m[15]);
        for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
        memcpy(m, t, sizeof m);
    }
    for (int i = 0; i < 8; i++) {
        out16[i] = s[i] ^ s[i + 8];
        out16[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void words_to_bytes(const uint32_t w[8], uint8_t out[32]) {
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)w[i];
        out[4 * i + 1] = (uint8_t)(w[i] >> 8);
        out[4 * i + 2] = (uint8_t)(w[i] >> 16);
        out[4 * i + 3] = (uint8_t)(w[i] >> 24);
    }
}

static void bytes_to_words(const uint8_t b[32], uint32_t w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}

/* Chaining value of one chunk (<= 1024 bytes) with the given chunk counter. */
static void chunk_cv(const uint8_t *data, size_t len, uint64_t counter, int root, uint32_t cv[8]) {
    size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    uint32_t h[8], out[16];
    memcpy(h, IV, sizeof h);
    for (size_t i = 0; i < nblocks; i++) {
        uint8_t blk[64];
        size_t off = i * 64;
        size_t bl = len - off < 64 ? len - off : 64;
        if (len == 0) bl = 0;
        memset(blk, 0, 64);
        if (bl) memcpy(blk, data + off, bl);
        uint32_t fl = (i == 0 ? CHUNK_START : 0) | (i == nblocks - 1 ? CHUNK_END : 0);
        if (root && i == nblocks - 1) fl |= ROOT;
        orc_blake3_compress(h, blk, counter, (uint32_t)bl, fl, out);
        memcpy(h, out, 32);
    }
    memcpy(cv, h, 32);
}

static void parent_words(const uint32_t l[8], const uint32_t r[8], uint32_t flags, uint32_t cv[8]) {
    uint8_t blk[64];
    uint32_t out[16];
    words_to_bytes(l, blk);
    words_to_bytes(r, blk + 32);
    orc_blake3_compress(IV, blk, 0, 64, PARENT | flags, out);
    memcpy(cv, out, 32);
}

/* BLAKE3 tree over chunks: the left subtree holds the largest power-of-two
 * number of chunks strictly less than the total. */
static void subtree_cv(const uint8_t *data, size_t len, uint64_t chunk_counter, int root,
                       uint32_t cv[8]) {
    size_t nchunks = len == 0 ? 1 : (len + CHUNK_LEN - 1) / CHUNK_LEN;
    if (nchunks == 1) {
        chunk_cv(data, len, chunk_counter, root, cv);
        return;
    }
    size_t left_chunks = 1;
    while (left_chunks * 2 < nchunks) left_chunks *= 2;
    size_t left_len = left_chunks * CHUNK_LEN;
    uint32_t l[8], r[8];
    subtree_cv(data, left_len, chunk_counter, 0, l);
    subtree_cv(data + left_len, len - left_len, chunk_counter + left_chunks, 0, r);
    parent_words(l, r, root ? ROOT : 0, cv);
}

void orc_blake3_hash(const uint8_t *data, size_t len, uint8_t out[32]) {
    uint32_t cv[8];
    subtree_cv(data, len, 0, 1, cv);
    words_to_bytes(cv, out);
}

/* Extended output (XOF): the root node's last compression, its input kept,
 * evaluated with ROOT and counter = output block index, 64 bytes a block. */
void orc_blake3_xof(const uint8_t *data, size_t len, uint8_t *out, size_t out_len) {
    uint32_t cv[8];
    uint8_t blk[64];
    uint32_t bl, fl;
    size_t nchunks = len == 0 ? 1 : (len + CHUNK_LEN - 1) / CHUNK_LEN;
    memset(blk, 0, sizeof blk);
    if (nchunks == 1) {
        /* single chunk: every block but the last chains from the IV */
        size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
        uint32_t o16[16];
        memcpy(cv, IV, sizeof cv);
        for (size_t i = 0; i + 1 < nblocks; i++) {
            orc_blake3_compress(cv, data + 64 * i, 0, 64, i == 0 ? CHUNK_START : 0, o16);
            memcpy(cv, o16, 32);
        }
        size_t off = 64 * (nblocks - 1);
        bl = (uint32_t)(len - off);
        if (bl) memcpy(blk, data + off, bl);
        fl = (nblocks == 1 ? CHUNK_START : 0) | CHUNK_END;
    } else {
        size_t left_chunks = 1;
        while (left_chunks * 2 < nchunks) left_chunks *= 2;
        size_t left_len = left_chunks * CHUNK_LEN;
        uint32_t l[8], r[8];
        subtree_cv(data, left_len, 0, 0, l);
        subtree_cv(data + left_len, len - left_len, left_chunks, 0, r);
        memcpy(cv, IV, sizeof cv);
        words_to_bytes(l, blk);
        words_to_bytes(r, blk + 32);
        bl = 64;
        fl = PARENT;
    }
    for (uint64_t b = 0; out_len; b++) {
        uint32_t o16[16];
        uint8_t bytes[64];
        orc_blake3_compress(cv, blk, b, bl, fl | ROOT, o16);
        for (int i = 0; i < 16; i++)
            for (int j = 0; j < 4; j++) bytes[4 * i + j] = (uint8_t)(o16[i] >> (8 * j));
        size_t take = out_len < 64 ? out_len : 64;
        memcpy(out, bytes, take);
        out += take;
        out_len -= take;
    }
}

void orc_blake3_leaf_cv(const uint8_t *data, size_t len, uint8_t out[32]) {
    uint32_t cv[8];
    subtree_cv(data, len, 0, 0, cv);
    words_to_bytes(cv, out);
}

void orc_blake3_parent_cv(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    uint32_t l[8], r[8], cv[8];
    bytes_to_words(left, l);
    bytes_to_words(right, r);
    parent_words(l, r, 0, cv);
    words_to_bytes(cv, out);
}

void orc_blake3_parent_root(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    uint32_t l[8], r[8], cv[8];
    bytes_to_words(left, l);
    bytes_to_words(right, r);
    parent_words(l, r, ROOT, cv);
    words_to_bytes(cv, out);
}

int orc_merkle_tree(const uint8_t *data, size_t data_len, size_t num_leaves, uint8_t *tree) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1)) || data_len % num_leaves) return -1;
    size_t leaf_size = data_len / num_leaves;
    for (size_t i = 0; i < num_leaves; i++)
        orc_blake3_leaf_cv(data + i * leaf_size, leaf_size, tree + 32 * i);
    size_t read_start = 0, read_len = num_leaves;
    while (read_len > 1) {
        size_t next = read_len >> 1;
        for (size_t i = 0; i < next; i++)
            orc_blake3_parent_cv(tree + 32 * (read_start + 2 * i), tree + 32 * (read_start + 2 * i + 1),
                                 tree + 32 * (read_start + read_len + i));
        read_start += read_len;
        read_len = next;
    }
    return 0;
}

int orc_merkle_proof(const uint8_t *tree, size_t num_leaves, size_t index, uint8_t *path_out) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1)) || index >= num_leaves) return -1;
    size_t level_start = 0, level_len = num_leaves, idx = index, k = 0;
    while (level_len > 1) {
        memcpy(path_out + 32 * k, tree + 32 * (level_start + (idx ^ 1)), 32);
        k++;
        level_start += level_len;
        level_len >>= 1;
        idx >>= 1;
    }
    return (int)k;
}

int orc_merkle_verify(const uint8_t root[32], const uint8_t leaf_hash[32], size_t index,
                      const uint8_t *path, size_t path_len) {
    uint8_t acc[32], tmp[32];
    memcpy(acc, leaf_hash, 32);
    size_t idx = index;
    for (size_t k = 0; k < path_len; k++) {
        if ((idx & 1) == 0)
            orc_blake3_parent_cv(acc, path + 32 * k, tmp);
        else
            orc_blake3_parent_cv(path + 32 * k, acc, tmp);
        memcpy(acc, tmp, 32);
        idx >>= 1;
    }
    return memcmp(acc, root, 32) == 0;
}

int orc_commit(const orc_f128 *msg, size_t n_msg, size_t num_ntts, size_t log_inv_rate,
               orc_f128 *codeword_out, uint8_t *tree_out, uint8_t root_out[32]) {
    if (num_ntts == 0 || n_msg % num_ntts || log_inv_rate < 1) return -1;
    size_t n_cw = n_msg << log_inv_rate;
    size_t n_pos = n_cw / num_ntts;
    size_t dim = 0;
    while (((size_t)1 << dim) < n_pos) dim++;
    int rc = orc_ntt_rs_encode_interleaved(dim, msg, n_msg, codeword_out, n_cw, num_ntts);
    if (rc) return rc;
    rc = orc_merkle_tree((const uint8_t *)codeword_out, n_cw * 16, n_pos, tree_out);
    if (rc) return rc;
    memcpy(root_out, tree_out + 32 * (2 * n_pos - 2), 32);
    return 0;
}
