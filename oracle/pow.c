/*
 * Proof-of-work grinding — CPU oracle restatement (TEST INFRASTRUCTURE ONLY;
 * only tests, the smoke check and the benchmark's CPU-baseline leg load it).
 *
 * Reference: CF/challenger.rs (CF = crates/flock-core/src/), BLAKE3
 * transcript:
 *   pre-image   state_digest (32 B) || nonce (u64 little-endian) || 24 zero
 *               bytes, one 64-byte block (blake3_pow_preimage, :549-554)
 *   hash        blake3::hash of the pre-image (pow_has_leading_zero_bits,
 *               :580-601)
 *   condition   at least `bits` leading zero bits, most significant bit of
 *               byte 0 first (has_leading_zero_bits, :558-571)
 *   grind       the smallest satisfying nonce, 0 when bits == 0 (grind_pow,
 *               :384-474: every search order returns the globally smallest)
 */
#include <string.h>

#include "flock_oracle.h"

void orc_pow_hash(const uint8_t digest[32], uint64_t nonce, uint8_t out[32])
{
    uint8_t pre[64];
    memset(pre, 0, sizeof pre);
    memcpy(pre, digest, 32);
    for (int i = 0; i < 8; ++i)
        pre[32 + i] = (uint8_t)(nonce >> (8 * i));
    orc_blake3_hash(pre, 64, out);
}

int orc_leading_zero_bits_ok(const uint8_t h[32], uint32_t bits)
{
    uint32_t full = bits / 8, extra = bits % 8;
    if (bits > 256)
        return 0;
    for (uint32_t i = 0; i < full; ++i)
        if (h[i])
            return 0;
    if (extra && (h[full] >> (8 - extra)))
        return 0;
    return 1;
}

int orc_pow_ok(const uint8_t digest[32], uint64_t nonce, uint32_t bits)
{
    uint8_t h[32];
    if (bits == 0)
        return nonce == 0; /* verify_pow's canonical zero-bit nonce (:502-517) */
    orc_pow_hash(digest, nonce, h);
    return orc_leading_zero_bits_ok(h, bits);
}

uint64_t orc_pow_grind(const uint8_t digest[32], uint32_t bits)
{
    if (bits == 0)
        return 0;
    uint8_t h[32];
    for (uint64_t n = 0;; ++n) {
        orc_pow_hash(digest, n, h);
        if (orc_leading_zero_bits_ok(h, bits))
            return n;
    }
}
