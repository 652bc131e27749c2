/*
 * Ligerito out-of-domain samples — CPU oracle (TEST INFRASTRUCTURE ONLY; only
 * tests, the smoke check and the benchmark's CPU-baseline leg load it).
 *
 * The OOD loop of the recursive prover (CF/pcs/ligerito.rs:10569-10590, CF =
 * crates/flock-core/src/) over the oracle's transcript byte log: per sample
 * z = sample_f128_vec(log_n), (u_0, u_2, y) = round_msg_and_eval_lsb(f, eq(z))
 * (introduce_new_with_eval, :9144-9154), observe y, u_0, u_2, beta =
 * sample_f128, then the glue b += beta eq(z), T += beta y (:9213-9256).
 */
#include <stdlib.h>

#include "flock_oracle.h"

int orc_ligerito_ood(orc_fs *t, const orc_f128 *f, orc_f128 *b, size_t log_n, size_t n_samples, orc_f128 *t_r,
                     orc_f128 *ood_values, orc_f128 *msgs)
{
    const size_t n = (size_t)1 << log_n;
    if (log_n < 1 || log_n > 40) return -1;
    orc_f128 *eq = malloc(n * sizeof *eq), *z = malloc(log_n * sizeof *z);
    int rc = (!eq || !z) ? -1 : 0;
    for (size_t s = 0; s < n_samples && !rc; ++s) {
        orc_f128 out[3], beta;
        if ((rc = orc_fs_sample_vec(t, log_n, z))) break;
        orc_build_eq(z, log_n, eq);
        orc_sc_round_msg_and_eval(f, eq, n, out);
        ood_values[s] = out[2];
        msgs[2 * s] = out[0];
        msgs[2 * s + 1] = out[1];
        if ((rc = orc_fs_f128(t, out[2]) || orc_fs_f128(t, out[0]) || orc_fs_f128(t, out[1]) ||
                  orc_fs_sample(t, &beta)))
            break;
        orc_sc_glue(b, eq, n, beta);
        orc_f128 p = orc_f128_mul(beta, out[2]);
        t_r->lo ^= p.lo;
        t_r->hi ^= p.hi;
    }
    free(eq);
    free(z);
    return rc;
}
