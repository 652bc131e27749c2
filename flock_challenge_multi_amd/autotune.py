"""Autotuner of the RS-encode pass schedules (the radix schedule of the additive NTT).

A batch of proofs commits one L0 codeword per size class: 2^(m-13+r) positions x
64 lanes for m = 14 + ceil(log2(n_blocks)) and the profile's rate 2^-r
(CP/r1cs_hashes/blake3.rs:4721-4775, CF/pcs/commit.rs:57-95).  For each such shape
`tune_encode` times the candidate pass schedules of the transform on random data
of that shape (HIP events on the launch stream), keeps the fastest and installs
it process-wide with flk_ntt_set_schedule.  A schedule only groups the same
layers into launches differently: every candidate computes the same codeword, so
proofs stay byte-identical (the batched tests compare every proof with its lone
proof).

The reference fixes its CPU schedule (a top/deep split sized for the caches,
CF/ntt/additive_ntt_f128.rs:4865-4900); here the search space is the pass sizes
of the device transform (1..3 layers per pass, the deepest pass 2 or 3) and their
order.
"""
import ctypes

import torch

from . import _native as nat
from . import hotpath as hp

LANES_LOG = 6  # the L0 commitment interleaves 64 columns (log_batch_size = 6)


def candidates(n_layers):
    """Pass schedules for n_layers layers (forward order, the deepest pass last):
    passes of 3 with the remainder as 2- or 1-layer passes, the short passes first
    or just before the deepest, the deepest pass 2 or 3 layers."""
    out = set()
    if n_layers <= 3:
        out.add((n_layers,))
    for deep in (2, 3):
        if deep >= n_layers:
            continue
        rest = n_layers - deep
        for n2 in range(4):
            r = rest - 2 * n2
            if r < 0 or r % 3:
                continue
            threes = [3] * (r // 3)
            out.add(tuple([2] * n2 + threes + [deep]))
            out.add(tuple(threes + [2] * n2 + [deep]))
        if rest % 3 == 1:
            out.add(tuple([1] + [3] * (rest // 3) + [deep]))
    return sorted(out)


def current(n_layers, lanes_log=LANES_LOG):
    """The schedule the library would use now for this shape."""
    ks = (ctypes.c_int * 64)()
    n = ctypes.c_size_t()
    nat.call("flk_ntt_get_schedule", n_layers, lanes_log, ks, 64, ctypes.byref(n))
    return tuple(ks[i] for i in range(n.value))


def install(n_layers, schedule, lanes_log=LANES_LOG):
    ks = (ctypes.c_int * max(1, len(schedule)))(*schedule)
    nat.call("flk_ntt_set_schedule", n_layers, lanes_log, ks, len(schedule))


def _time_encode(log_d, rate, msg, cw, reps):
    st = torch.cuda.current_stream()
    sptr = ctypes.c_void_p(st.cuda_stream)

    def enc():
        nat.call("flk_ntt_rs_encode_interleaved", log_d, ctypes.c_void_p(msg.data_ptr()), msg.shape[0],
                 ctypes.c_void_p(cw.data_ptr()), cw.shape[0], 1 << LANES_LOG, sptr)

    enc()  # twiddles cached, first-launch costs out of the timing
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(reps):
        enc()
    b.record(st)
    b.synchronize()
    return a.elapsed_time(b) / reps


def tune_encode(log_d, rate, reps=5, install_best=True):
    """Time every candidate schedule of the rate-2^-rate encode onto 2^log_d
    positions x 64 lanes; install the fastest.  Returns {schedule: ms} and the
    choice.  The encode runs log_d - rate layers (the top `rate` layers only
    replicate the message)."""
    n_layers = log_d - rate
    g = torch.Generator(device="cuda")
    g.manual_seed(log_d * 131 + rate)
    msg = torch.randint(-(2**63), 2**63 - 1, ((1 << (log_d - rate + LANES_LOG)), 2), dtype=torch.int64,
                        device="cuda", generator=g)
    cw = hp.f128_empty(1 << (log_d + LANES_LOG))
    install(n_layers, ())  # a schedule installed earlier is not the library's rule
    before = current(n_layers)
    times = {}
    try:
        for sched in candidates(n_layers):
            install(n_layers, sched)
            times[sched] = _time_encode(log_d, rate, msg, cw, reps)
    finally:
        install(n_layers, ())  # back to the default rule
    best = min(times, key=times.get)
    if install_best:
        install(n_layers, best)
    del msg, cw
    return {"log_d": log_d, "rate": rate, "default": list(before), "default_ms": times.get(before),
            "best": list(best), "best_ms": times[best],
            "tried": {",".join(map(str, k)): v for k, v in sorted(times.items(), key=lambda kv: kv[1])}}


def tune_batch(m_values, rate=1, reps=5):
    """Tune the L0 encode of every size class m of a batch (log_d = m - 13 + rate)."""
    return [tune_encode(m - 13 + rate, rate, reps) for m in sorted(set(m_values))]
