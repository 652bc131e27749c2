/*
 * Oracle (test infrastructure): the commit of orc_commit (blake3_merkle.c)
 * computed by n_threads POSIX threads, for the CPU baseline of the benchmark.
 * Same arithmetic, same order of operations per element: the RS encode is the
 * zero-padded full forward transform (CF/ntt/additive_ntt_f128.rs:4766-4795,
 * :2785-2830), each layer's butterflies split into contiguous ranges of
 * (block, row) pairs; the Merkle tree (CF/merkle.rs:753-815) hashes leaves and
 * then each level's parents in contiguous ranges.  Threads meet between
 * layers and levels, so every butterfly reads the previous layer's values.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static orc_f128 f_add(orc_f128 a, orc_f128 b) {
    orc_f128 r = {a.lo ^ b.lo, a.hi ^ b.hi};
    return r;
}

typedef void (*range_fn)(void *ctx, size_t lo, size_t hi);

struct range_job {
    range_fn fn;
    void *ctx;
    size_t lo, hi;
};

static void *range_main(void *p) {
    struct range_job *j = (struct range_job *)p;
    j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}

/* Run fn over [0, n) in n_threads contiguous ranges; the calling thread takes the first. */
int orc_parallel_for(size_t n, size_t n_threads, range_fn fn, void *ctx) {
    if (n_threads < 1) n_threads = 1;
    if (n_threads > n) n_threads = n ? n : 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    struct range_job jobs[256];
    size_t per = (n + n_threads - 1) / n_threads;
    size_t started = 0;
    for (size_t t = 1; t < n_threads; t++) {
        jobs[t].fn = fn;
        jobs[t].ctx = ctx;
        jobs[t].lo = t * per < n ? t * per : n;
        jobs[t].hi = (t + 1) * per < n ? (t + 1) * per : n;
        if (pthread_create(&th[t], NULL, range_main, &jobs[t]) != 0) break;
        started = t;
    }
    fn(ctx, 0, per < n ? per : n);
    for (size_t t = 1; t <= started; t++) pthread_join(th[t], NULL);
    /* ranges whose thread could not be started run here */
    for (size_t t = started + 1; t < n_threads; t++) fn(ctx, jobs[t].lo, jobs[t].hi);
    return 0;
}

struct layer_ctx {
    orc_f128 *data;
    const orc_f128 *tw;
    size_t num_blocks, half, num_ntts;
};

static void layer_range(void *p, size_t lo, size_t hi) {
    struct layer_ctx *c = (struct layer_ctx *)p;
    for (size_t j = lo; j < hi; j++) {
        size_t block = j / c->half, row = j % c->half;
        orc_f128 t = c->tw[c->num_blocks - 1 + block];
        size_t off_top = (block * 2 * c->half + row) * c->num_ntts;
        size_t off_bot = off_top + c->half * c->num_ntts;
        for (size_t lane = 0; lane < c->num_ntts; lane++) {
            orc_f128 v = c->data[off_bot + lane];
            orc_f128 u = f_add(c->data[off_top + lane], orc_f128_mul(v, t));
            c->data[off_top + lane] = u;
            c->data[off_bot + lane] = f_add(v, u);
        }
    }
}

struct leaf_ctx {
    const uint8_t *data;
    size_t leaf_size;
    uint8_t *tree;
};

static void leaf_range(void *p, size_t lo, size_t hi) {
    struct leaf_ctx *c = (struct leaf_ctx *)p;
    for (size_t i = lo; i < hi; i++) orc_blake3_leaf_cv(c->data + i * c->leaf_size, c->leaf_size, c->tree + 32 * i);
}

struct level_ctx {
    uint8_t *tree;
    size_t read_start, read_len;
};

static void level_range(void *p, size_t lo, size_t hi) {
    struct level_ctx *c = (struct level_ctx *)p;
    for (size_t i = lo; i < hi; i++)
        orc_blake3_parent_cv(c->tree + 32 * (c->read_start + 2 * i), c->tree + 32 * (c->read_start + 2 * i + 1),
                             c->tree + 32 * (c->read_start + c->read_len + i));
}

int orc_commit_mt(const orc_f128 *msg, size_t n_msg, size_t num_ntts, size_t log_inv_rate,
                  orc_f128 *codeword_out, uint8_t *tree_out, uint8_t root_out[32], size_t n_threads) {
    if (num_ntts == 0 || (num_ntts & (num_ntts - 1)) || n_msg % num_ntts || log_inv_rate < 1) return -1;
    size_t n_cw = n_msg << log_inv_rate;
    size_t n_pos = n_cw / num_ntts;
    size_t log_d = 0;
    while (((size_t)1 << log_d) < n_pos) log_d++;
    if (((size_t)1 << log_d) != n_pos) return -1;
    orc_f128 *tw = (orc_f128 *)malloc((((size_t)1 << log_d) - 1) * sizeof(orc_f128));
    if (!tw) return -2;
    orc_ntt_twiddles(log_d, tw);
    memcpy(codeword_out, msg, n_msg * sizeof(orc_f128));
    memset(codeword_out + n_msg, 0, (n_cw - n_msg) * sizeof(orc_f128));
    for (size_t layer = 0; layer < log_d; layer++) {
        struct layer_ctx c = {codeword_out, tw, (size_t)1 << layer, (size_t)1 << (log_d - layer - 1), num_ntts};
        orc_parallel_for(n_pos / 2, n_threads, layer_range, &c);
    }
    free(tw);
    struct leaf_ctx lc = {(const uint8_t *)codeword_out, num_ntts * 16, tree_out};
    orc_parallel_for(n_pos, n_threads, leaf_range, &lc);
    size_t read_start = 0, read_len = n_pos;
    while (read_len > 1) {
        struct level_ctx vc = {tree_out, read_start, read_len};
        orc_parallel_for(read_len / 2, n_threads, level_range, &vc);
        read_start += read_len;
        read_len >>= 1;
    }
    memcpy(root_out, tree_out + 32 * (2 * n_pos - 2), 32);
    return 0;
}
