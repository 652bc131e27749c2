// Ligerito out-of-domain samples at a recursive level (CF = crates/flock-core/src/):
//
//   flk_ligerito_ood   the OOD loop of the recursive prover (CF/pcs/ligerito.rs:10569-10590):
//                      per sample, z = sample_f128_vec(log_n); (u_0, u_2, y) =
//                      round_msg_and_eval_lsb(f, eq(z)) (introduce_new_with_eval, :9144-9154);
//                      observe y, u_0, u_2; beta = sample_f128; glue: b += beta eq(z),
//                      T += beta y (:9213-9256)
//
// eq(z) is built on the device into stream-ordered scratch, the message and
// evaluation are one reduction over (f, eq(z)), and the glue one pass over b;
// the three values come back to the host for the transcript.
#include <hip/hip_runtime.h>

#include "internal.hpp"

using namespace flk;

extern "C" {

int flk_ligerito_ood(flk_fs* fs, const flk_f128* f, flk_f128* b, size_t log_n, size_t n_samples, flk_f128* t_r,
                     flk_f128* ood_values_out, flk_f128* msgs_out, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!fs || !f || !b || !t_r || (n_samples && (!ood_values_out || !msgs_out)))
    return set_error(FLK_E_SHAPE, "ligerito_ood: null pointer");
  if (log_n < 1 || log_n > 34) return set_error(FLK_E_SHAPE, "ligerito_ood: log_n=%zu (1..34)", log_n);
  if (n_samples > 64) return set_error(FLK_E_SHAPE, "ligerito_ood: %zu samples (at most 64)", n_samples);
  if (n_samples == 0) return FLK_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t n = size_t(1) << log_n;
  uint4* ws = nullptr;  // eq(z) then the three reduction outputs
  hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&ws), (n + 4) * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "ligerito_ood: workspace allocation failed");
  }
  flk_f128* eq = reinterpret_cast<flk_f128*>(ws);
  flk_f128* red = reinterpret_cast<flk_f128*>(ws + n);
  flk_f128 z[34], out3[3];
  for (size_t s = 0; s < n_samples && rc == FLK_OK; ++s) {
    if ((rc = flk_fs_sample_f128_vec(fs, log_n, z))) break;
    if ((rc = flk_build_eq(z, log_n, eq, st))) break;
    if ((rc = flk_sc_round_msg_and_eval(f, eq, n, red, st))) break;
    e = read_to_host(out3, red, sizeof out3, st);
    if (e != hipSuccess) {
      rc = device_failure(e, "ligerito_ood: readback");
      break;
    }
    const flk_f128 y = out3[2];
    ood_values_out[s] = y;
    msgs_out[2 * s] = out3[0];
    msgs_out[2 * s + 1] = out3[1];
    flk_f128 beta;
    if ((rc = flk_fs_observe_f128(fs, y)) || (rc = flk_fs_observe_f128(fs, out3[0])) ||
        (rc = flk_fs_observe_f128(fs, out3[1])) || (rc = flk_fs_sample_f128(fs, &beta)))
      break;
    if ((rc = flk_sc_glue(b, eq, n, beta, st))) break;
    const flk_f128 p = host_mul(beta, y);
    t_r->lo ^= p.lo;
    t_r->hi ^= p.hi;
  }
  e = flk::stream_free(ws, st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (rc == FLK_OK) rc = device_failure(e, "ligerito_ood");
  }
  return rc;
}

}  // extern "C"
