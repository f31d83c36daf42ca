"""The validation harness's host side (nbvh_amd/validation.py): the reference's own configuration
file read into harness settings, and the OpenEXR files it writes read back."""
import json
import os

import numpy as np

from conftest import GOLDEN

CHESS = os.path.join(GOLDEN, "chess_nodes_32940_hash_18_lvl_8_neural.json")


def test_reference_configuration_settings(nb):
    from nbvh_amd import validation as V
    s = V.settings(json.load(open(CHESS)))
    assert s["mode"] == "Full" and s["base"] == "nodes_32940_hash_18_lvl_8_tsteps_10000" and s["variant"] == "neural"
    assert s["steps"] == 10000 and s["spp"] == 4096 and s["max_depth"] == 3 and s["rr_start_depth"] == 2
    L = s["learner"]
    assert L["batch_size"] == 262144 and L["max_training_steps"] == 10000 and L["tree_cut_learning"] == 1
    assert (L["splitting_interval"], L["last_split_iteration"], L["start_split_num"]) == (40, 3000, 1)
    assert abs(L["split_scaling"] - 2.5) < 1e-7 and abs(L["min_rr_survival_prob"] - 0.005) < 1e-9
    assert L["num_lods_learned"] == 1 and L["max_training_bounces"] == 1 and L["training_node_survival"] == 1
    g = s["network"].grid
    assert (g.n_levels, g.n_features_per_level, g.log2_hashmap_size, g.base_resolution) == (8, 4, 18, 8)
    assert abs(g.per_level_scale - 2.0) < 1e-7 and s["network"].n_points == 3 and s["network"].segment_distribution == 1
    t = s["trainer"]
    assert (t["decay_start"], t["decay_interval"]) == (10000, 2000) and abs(t["learning_rate"] - 0.01) < 1e-9
    assert s["fb_size"] == (1920, 1080) and s["primary_threshold"] == 0.5 and s["secondary_threshold"] == 0.5
    assert s["sky"] is False


def test_exr_round_trip(tmp_path):
    from nbvh_amd import validation as V
    img = np.random.default_rng(0).uniform(0, 3, (13, 21, 3)).astype(np.float32)
    p = tmp_path / "x.exr"
    V.write_exr(str(p), img)
    data = p.read_bytes()
    assert data[:4] == bytes([0x76, 0x2F, 0x31, 0x01]) and data[4] == 2  # OpenEXR magic, version 2
    assert np.array_equal(V.read_exr(str(p)), img)
