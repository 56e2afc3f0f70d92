"""CPU tests of the autotuner's candidate set and of the determinism of its
choice given a table (no device needed)."""
from mustard_amd import autotune


def test_candidates():
    assert autotune.candidates(37) == [1]
    assert autotune.candidates(512) == [1, 2, 4, 8, 16]
    assert autotune.candidates(1000) == [1, 2, 4, 5, 8, 10, 20, 25]
    for n in (96, 300, 1024, 24000):
        for t in autotune.candidates(n):
            assert n % t == 0 and (t == 1 or 32 <= n // t <= 1024)


def test_choose_is_deterministic_and_nearest():
    table = {"chol": {"256": 2, "512": 4}, "lu": {"1024": 8}}
    # order 512 tabulated: b0 = 128 -> t = 4
    assert autotune.choose("chol", 512, table) == 4
    # order 500 -> nearest 512, b0 = 128: candidates of 500 are 1, 2, 4, 5, 10
    # with b = 500, 250, 125, 100, 50 -> 125 is nearest
    assert autotune.choose("chol", 500, table) == 4
    # ties between tabulated orders take the smaller one (384 -> 256, b0 = 128)
    assert autotune.choose("chol", 384, table) == 3
    # a prime order has only t = 1
    assert autotune.choose("lu", 997, table) == 1
    # missing kind -> 1
    assert autotune.choose("qr", 512, table) == 1
    for _ in range(3):
        assert autotune.choose("lu", 768, table) == autotune.choose("lu", 768, dict(table))
