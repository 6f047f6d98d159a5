"""Tuning parameters and the autotuner: every setting gives the oracle's
answer; the tuner leaves a valid setting on the graph and caches it."""
import json

import numpy as np
import pytest

import gt_hip
import oracle
from gt_hip import autotune

pytestmark = pytest.mark.gpu


def _graph(scale=13, seed=17):
    n = 1 << scale
    s, t = oracle.gen_rmat(scale, 16 * n, 0.57, 0.19, 0.19, seed)
    return n, s, t


def test_tuning_round_trip_and_defaults(gpu):
    n, s, t = _graph(8)
    dev = gt_hip.DeviceGraph.from_edges(n, s, t)
    d = dev.tuning()
    assert d == {"sweep_blocks_per_cu": 3, "tile_rows": 256, "short_row": 64,
                 "mm_tile_nnz": 1024, "mm_blocks_per_cu": 8, "sweep_hot_kb": 128,
                 "sliced_min_len": -1, "sliced_share": 0, "block_tile_nnz": 131072,
                 "block_long_min": 2048, "block_per_cu": 2, "block_f64": 1}
    got = dev.set_tuning(tile_rows=1000, short_row=8, sweep_blocks_per_cu=99, mm_tile_nnz=10)
    assert got["tile_rows"] == 256 and got["short_row"] == 8
    assert got["sweep_blocks_per_cu"] == 8 and got["mm_tile_nnz"] == 256
    assert dev.set_tuning(tile_rows=0, short_row=0)["tile_rows"] == 256
    assert dev.set_tuning(sweep_hot_kb=500)["sweep_hot_kb"] == 128
    assert dev.set_tuning(sweep_hot_kb=-7)["sweep_hot_kb"] == -1
    assert dev.set_tuning(sliced_min_len=1)["sliced_min_len"] == 2
    assert dev.set_tuning(sliced_min_len=5000, sliced_share=150)["sliced_min_len"] == 1024
    assert dev.tuning()["sliced_share"] == 99
    got = dev.set_tuning(block_tile_nnz=10, block_long_min=1 << 20)
    assert got["block_tile_nnz"] == 1024 and got["block_long_min"] == 65536
    got = dev.set_tuning(block_tile_nnz=-5, block_long_min=3)
    assert got["block_tile_nnz"] == -1 and got["block_long_min"] == 16
    assert dev.set_tuning(block_per_cu=9)["block_per_cu"] == 3
    assert dev.set_tuning(block_f64=-7)["block_f64"] == -1
    assert dev.set_tuning(block_f64=0)["block_f64"] == 1
    assert dev.set_tuning(sliced_min_len=0, sliced_share=-3)["sliced_min_len"] == -1
    assert dev.tuning()["sliced_share"] == 0
    with pytest.raises(ValueError):
        dev.set_tuning(bogus=1)


@pytest.mark.parametrize("tune", [
    {"tile_rows": 16, "short_row": 1, "sweep_blocks_per_cu": 1},
    {"tile_rows": 64, "short_row": 1024, "sweep_blocks_per_cu": 8},
    {"tile_rows": 128, "short_row": 16, "sweep_blocks_per_cu": 4},
    {"sweep_hot_kb": -1},
    {"sweep_hot_kb": 1, "tile_rows": 64},
    {"sweep_hot_kb": 2, "sliced_min_len": 64, "sliced_share": 25},
])
def test_every_setting_matches_the_oracle(gpu, tune):
    n, s, t = _graph()
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-10, grain=64)
    g.device().set_tuning(**tune)
    r = gt_hip.DenseProp.new(g.graph_id)
    it = gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85,
                         gt_hip.PowerIteration(1e-10, 0, gt_hip.Plan(n, 64)), r)
    assert abs(it - it_o) <= 1
    assert np.abs(r.data - want).sum() <= 1e-12
    x = gt_hip.DenseProp.new(g.graph_id)
    eig, ite = gt_hip.eigenvector(g, gt_hip.Unity.NEW,
                                  gt_hip.PowerIteration(1e-10, 0, gt_hip.Plan(n, 64)), x)
    xo, eo, ito = oracle.Adj(n, s, t).eigenvector(epsilon=1e-10, grain=64)
    assert abs(ite - ito) <= 1 and abs(eig - eo) <= 1e-9 * abs(eo)


def test_autotune_picks_a_setting_and_caches_it(gpu, tmp_path):
    n, s, t = _graph(14)
    dev = gt_hip.DeviceGraph.from_edges(n, s, t)
    cache = tmp_path / "tuning.json"
    rec = autotune.autotune(dev, "pagerank", cache=cache, sweeps=3, reps=2)
    assert rec["tuning"] == {k: dev.tuning()[k] for k in autotune.SWEEP_SPACE}
    assert rec["sweep_us"] <= rec["default_sweep_us"] * 1.05
    assert len(rec["trials"]) >= 1 + sum(len(v) - 1 for v in autotune.SWEEP_SPACE.values())
    data = json.loads(cache.read_text())
    assert len(data) == 1
    dev.set_tuning(**{k: 0 for k in autotune.SWEEP_SPACE})
    assert autotune.apply_cached(dev, "pagerank", cache=cache) == rec["tuning"]
    rb = autotune.autotune(dev, "batched", cache=cache, sweeps=2, reps=1)
    assert rb["tuning"]["mm_tile_nnz"] in autotune.BATCHED_SPACE["mm_tile_nnz"]
    assert len(json.loads(cache.read_text())) == 2
