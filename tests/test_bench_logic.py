"""Host logic of bench.py's replica ensemble (no GPU needed): the 512-replica configuration and
its split into batches by model and size."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _all_idx(groups):
    return np.sort(np.concatenate([g["idx"] for g in groups.values()]))


def test_ensemble_config_covers_every_replica_once():
    groups = bench.ensemble_config()
    assert set(groups) == {3, 4}
    np.testing.assert_array_equal(_all_idx(groups), np.arange(512))
    for model, g in groups.items():
        assert len(g["states"]) == len(g["ns"]) == len(g["nats"]) == len(g["seeds"])
        for n, st in zip(g["ns"], g["states"]):
            assert st.size == n * model
        assert np.all((g["ns"] >= 2048) & (g["ns"] <= 65536))


def test_split_by_size_partitions_each_model():
    groups = bench.ensemble_config()
    for n_split in (12288, 24576, 65536 + 1):
        split = bench.split_by_size(groups, n_split)
        # every replica in exactly one batch, of its own model, on the side of the threshold it falls
        np.testing.assert_array_equal(_all_idx(split), np.arange(512))
        for key, g in split.items():
            assert g["model"] in (3, 4) and key.startswith(f"model{g['model']}")
            assert g["ns"].size > 0
            big = g["ns"] >= n_split
            assert big.all() if key.endswith("_large") else not big.any()
            src = groups[g["model"]]
            pos = {r: k for k, r in enumerate(src["idx"])}
            for k, r in enumerate(g["idx"]):
                j = pos[r]  # the same replica's data, carried unchanged
                assert g["ns"][k] == src["ns"][j] and g["seeds"][k] == src["seeds"][j]
                assert g["sigma"][k] == src["sigma"][j] and g["D"][k] == src["D"][j] and g["xi"][k] == src["xi"][j]
                assert g["states"][k] is src["states"][j] and g["nats"][k] is src["nats"][j]


def test_split_zero_keeps_one_batch_per_model():
    groups = bench.ensemble_config()
    split = bench.split_by_size(groups, 0)
    assert set(split) == {"model3", "model4"}
    for model in (3, 4):
        np.testing.assert_array_equal(split[f"model{model}"]["idx"], groups[model]["idx"])
