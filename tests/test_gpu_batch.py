"""GPU tests of batched uneven factorizations (mu_record_batch: many
independent problems in one DAG, one graph launch) and of the autotuner's
tile-size sweep against the oracle."""
import numpy as np
import pytest

from tests.util import EPS

pytestmark = pytest.mark.gpu


def _rel(got, ref, lower_only):
    if lower_only:
        m = np.tril(np.ones(ref.shape, dtype=bool))
        got, ref = got[m], ref[m]
    return np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref))))


def _check(oracle, kind, A, F, t):
    ref = np.array(A, order="F", copy=True)
    if kind == "lu":
        assert oracle.tiled_lu(ref, t) == 0
    else:
        assert oracle.tiled_cholesky(ref, t) == 0
    assert _rel(F, ref, kind == "chol") <= 1e-10


def test_batch_uneven_vs_oracle(torch_gpu, oracle):
    from mustard_amd import autotune
    from mustard_amd.batch import Batch, make_problems, unpack
    rng = np.random.default_rng(5)
    kinds = ["chol"] * 96 + ["lu"] * 24
    ns = list(rng.integers(32, 513, 96)) + list(rng.integers(64, 1025, 24))
    table = {"chol": {"128": 1, "256": 2, "512": 4}, "lu": {"512": 4, "1024": 8}}
    tiles = [autotune.choose(k, n, table) for k, n in zip(kinds, ns)]
    host = make_problems(kinds, ns, 17)
    b = Batch(kinds, ns, tiles)
    b.load(host)
    b.launch()
    torch_gpu.cuda.synchronize()
    assert not b.infos().any()
    for p in range(len(ns)):
        _check(oracle, kinds[p], unpack(host, ns, p), b.factor(p), tiles[p])
    assert b.exec.stats()["levels"] < 200  # one graph of level kernels for all problems
    b.close()


def test_batch_graph_replays(torch_gpu):
    """Replays of the batch graph, with eager device work and host
    synchronisation between them, give bit-identical factors."""
    from mustard_amd.batch import Batch, make_problems
    kinds = ["chol", "lu"] * 20
    ns = [64 + 24 * (p % 17) for p in range(40)]
    ns = [n - n % 4 for n in ns]
    host = make_problems(kinds, ns, 3)
    b = Batch(kinds, ns, [4] * 40)
    outs = []
    for r in range(3):
        b.load(host)
        scratch = torch_gpu.randn(1 << 20, device="cuda", dtype=torch_gpu.float64)
        scratch.mul_(2.0).sum().item()  # eager work between replays
        b.launch()
        torch_gpu.cuda.synchronize()
        assert not b.infos().any()
        outs.append(b.buf.clone())
    assert all(bool((o == outs[0]).all()) for o in outs)
    b.close()


@pytest.mark.parametrize("kind,n", [("chol", 512), ("lu", 512), ("chol", 300), ("lu", 1000)])
def test_tile_size_sweep_vs_oracle(torch_gpu, oracle, kind, n):
    """Every autotuner candidate factors to the oracle's tolerance."""
    from mustard_amd import autotune
    from mustard_amd.batch import Batch, make_problems, unpack
    for t in autotune.candidates(n):
        host = make_problems([kind] * 2, [n] * 2, t)
        b = Batch([kind] * 2, [n] * 2, [t] * 2)
        b.load(host)
        b.launch()
        torch_gpu.cuda.synchronize()
        assert not b.infos().any()
        for p in range(2):
            _check(oracle, kind, unpack(host, [n] * 2, p), b.factor(p), t)
        b.close()


def test_autotune_measures_and_chooses(torch_gpu, tmp_path):
    from mustard_amd import autotune
    table = autotune.tune(kinds=("chol",), orders=(64, 256), path=str(tmp_path / "t.json"),
                          log=None)
    assert set(table["chol"]) == {"64", "256"}
    for k, t in table["chol"].items():
        assert t in autotune.candidates(int(k))
    again = autotune.load(str(tmp_path / "t.json"))
    assert autotune.choose("chol", 250, again) == autotune.choose("chol", 250, table)
