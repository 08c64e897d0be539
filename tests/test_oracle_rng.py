"""Pins of the counter-based draws behind the thermostats (CPU only)."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O
from xla_cco.dynamics import rng as R

# Published known-answer vectors of Philox-4x32-10 (Random123): counter, key, result.
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    (
        (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
        (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1),
    ),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want) -> None:
    assert O.philox4x32_10(ctr, key) == want
    got = R.philox(np.array([ctr], dtype=np.uint64), np.array(key, dtype=np.uint32))[0]
    assert tuple(int(x) for x in got) == want


def test_host_draws_equal_the_oracle() -> None:
    rng = np.random.default_rng(3)
    for _ in range(20):
        key = rng.integers(0, 2**32, size=2, dtype=np.uint64).astype(np.uint32)
        kt = (int(key[0]), int(key[1]))
        assert [tuple(int(x) for x in k) for k in R.split(key, 4)] == O.rng_split(kt, 4)
        assert np.array_equal(R.normals(key, 7), O.rng_normals(kt, 7))
        assert R.uniform(key) == O.rng_uniform(kt)
        assert R.randint(key, 37) == O.rng_randint(kt, 37)


def test_draw_statistics() -> None:
    z = R.normals(R.prng_key(11), 200_000)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3
    keys = R.split(R.prng_key(12), 20_000)
    u = np.array([R.uniform(k) for k in keys])
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    ints = np.array([R.randint(k, 5) for k in keys])
    assert set(np.unique(ints)) == {0, 1, 2, 3, 4}
    assert stats.chisquare(np.bincount(ints, minlength=5)).pvalue > 1e-3
    # split keys are distinct and differ from their parent
    k = R.prng_key(13)
    ks = R.split(k, 8)
    assert len({tuple(x) for x in ks} | {tuple(k)}) == 9


def test_keys_are_checked() -> None:
    with pytest.raises(TypeError):
        R.require_key(np.array([1, 2], dtype=np.int64))
    with pytest.raises(ValueError):
        R.prng_key(-1)
