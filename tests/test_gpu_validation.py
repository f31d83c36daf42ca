"""The validation harness end to end on the GPU: a Full configuration (tree-cut learning and
training, then rendering and FLIP against the triangle rendering) and an Inference configuration
that reloads what the first one saved and renders the same image bit for bit."""
import csv
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_full_then_inference(nb, tmp_path):
    from nbvh_amd import validation as V
    base = {"cuda_scene": {"fb_size": [96, 64], "scene_constants": {"skylike_background": True}},
            "neural_bvh_renderer": {
                "validation_mode": "Full", "base_config_name": "ico", "variant_config_name": "neural",
                "batch_size": 16384, "max_training_steps": 120, "max_accumulated_spp": 3,
                "tree_cut_learning": True, "bvh_split_scheduler": {"splitting_interval": 10, "last_split_iteration": 100},
                "neural_module": {"max_depth": 3, "segment_encoding": {"hashgrid_options": {
                    "n_levels": 8, "n_features_per_level": 4, "log2_hashmap_size": 14, "base_resolution": 8,
                    "per_level_scale": 2.0}}}}}
    full = tmp_path / "full.json"
    full.write_text(json.dumps(base))
    base["neural_bvh_renderer"]["validation_mode"] = "Inference"
    base["neural_bvh_renderer"]["variant_config_name"] = "again"
    inf = tmp_path / "inference.json"
    inf.write_text(json.dumps(base))
    out = tmp_path / "out"
    res = V.run([str(full), str(inf)], str(out), scene_spec="icosphere:4:3:0.1", log=lambda m: print(m, flush=True))
    tr = res[0]["training"]
    assert tr["steps"] == 120 and tr["n_nodes"] > 1 and tr["sample_failures"] == 0
    assert tr["loss_last"] < tr["loss_first"]
    rows = list(csv.reader(open(out / "results" / "ico_neural_learning_spp_0.csv")))
    assert rows[0] == ["frame index", "frame time (ms)", " loss"] and len(rows) == 121
    assert nb.nbvh_parse(open(out / "nbvhs" / "ico.nbvh", "rb").read()).shape[0] == tr["n_nodes"]
    for r in res:
        f = r["inference"]["flip"]
        assert 0.0 < f < 1.0
    a = V.read_exr(str(out / "results" / "ico_neural_spp_3.exr"))
    b = V.read_exr(str(out / "results" / "ico_again_spp_3.exr"))
    assert a.shape == (64, 96, 3) and np.array_equal(a, b)  # reloaded N-BVH and network render the same
    assert res[0]["inference"]["flip"] == res[1]["inference"]["flip"]
    ref = V.read_exr(str(out / "results" / "ico_reference_spp_3.exr"))
    assert ref.max() > 0.1
    assert os.path.exists(out / "results" / "ico_neural_inference_spp_3.csv")
