"""Host logic of the recursive Ligerito prover (flock_challenge_multi_amd.ligerito)
that runs without a GPU: ceil_log2 (CF/pcs/ligerito.rs:1338-1344) and the
ProverConfig shape checks the prover makes before touching the device."""
import pytest

from flock_challenge_multi_amd import ligerito as lig


def test_ceil_log2():
    assert [lig.ceil_log2(n) for n in (0, 1, 2, 3, 4, 5, 8, 9, 218, 256, 257)] == [0, 0, 1, 2, 2, 3, 3, 4, 8, 8, 9]


def cfg(**kw):
    base = dict(initial_k=3, recursive_ks=[2, 2], log_inv_rates=[1, 2, 2], queries=[8, 6, 5],
                grinding_bits=[2, 2, 2], fold_grinding_bits=[3, 2, 1], ood_samples=[0, 1, 1])
    base.update(kw)
    return lig.ProverConfig(**base)


def test_config_checks():
    cfg().check(12)
    with pytest.raises(ValueError):
        cfg(recursive_ks=[]).check(12)
    with pytest.raises(ValueError):
        cfg(queries=[8, 6]).check(12)
    with pytest.raises(ValueError):
        cfg(ood_samples=[1, 1, 1]).check(12)  # L0 is bound by the opening's own claim
    with pytest.raises(ValueError):
        cfg().check(6)  # 3 + 2 + 2 folds over 6 variables
    # the m = 32 Fast ladder
    lig.ProverConfig(initial_k=6, recursive_ks=[3] * 5, log_inv_rates=[1, 2, 3, 4, 5, 6],
                     queries=[218, 106, 71, 53, 43, 36], grinding_bits=[0] * 6,
                     fold_grinding_bits=[19, 14, 11, 8, 6, 4], ood_samples=[0, 1, 1, 1, 1, 1]).check(25)
