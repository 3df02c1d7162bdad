"""Launch-shape knobs and the autotuner on the GPU (hpvcd_gpu_set_tuning,
hpvcd_tune, HPVCD_TUNING_FILE).

Every value of every knob must give the same output samples as the default
shapes (bit-exact), in batched and graph mode (a graph captured with other
shapes is captured again); the tuner's choice too; and a tuning cache file
named by HPVCD_TUNING_FILE is applied when a GPU is opened.
"""
import os

import numpy as np
import pytest

import hpvcd_amd as H
from hpvcd_amd import _native as N

pytestmark = pytest.mark.gpu

VALUES = {"residual_block": [64, 128], "deblock_block": [64, 128], "rgb_rows": [1, 2], "intra_per_cu": [1, 2],
          "sao_parts": [1, 2]}

SETS = {
    "grid3x2": dict(workload="grid4032", grid_cols=3, grid_rows=2, out_width=1400, out_height=1000, base_seed=501),
    "rext444_10": dict(workload="rext4096", grid_cols=2, grid_rows=1, out_width=1024, out_height=512,
                       base_seed=503),
    "inter_gop8": dict(workload="single512", tile_width=256, tile_height=128, grid_cols=1, grid_rows=9,
                       num_outputs=9, out_width=256, out_height=128, inter=1, weighted=1, gop=8, base_seed=505),
}


def _variants():
    d = H.default_tuning()
    for k, vals in VALUES.items():
        for v in vals:
            if k == "residual_block":
                for i in range(4):
                    rb = list(d["residual_block"])
                    rb[i] = v
                    yield f"{k}[{i}]={v}", dict(d, residual_block=rb)
            else:
                yield f"{k}={v}", dict(d, **{k: v})


@pytest.mark.parametrize("name", list(SETS))
def test_every_knob_value_gives_the_same_samples(name):
    rs = H.RecordSet.generate(**SETS[name])
    with H.Gpu(0) as g:
        g.tuning = H.default_tuning()
        with g.job(rs) as job:
            job.run(N.MODE_GRAPH)
            want = job.download()
            for label, t in _variants():
                print(f"tuning {name} {label}", flush=True)
                g.tuning = t
                for mode in (N.MODE_BATCHED, N.MODE_GRAPH):
                    job.run(mode)
                    got = job.download()
                    assert all(np.array_equal(a, b) for a, b in zip(got, want)), (label, mode)


def test_tuner_choice_and_cache_file(tmp_path):
    rs = H.RecordSet.generate(**SETS["grid3x2"])
    f = str(tmp_path / "tune.txt")
    with H.Gpu(0) as g:
        g.tuning = H.default_tuning()
        with g.job(rs) as job:
            job.run()
            want = job.download()
            print("tuning: autotuner search", flush=True)
            best, best_ms, default_ms = job.tune(iters=3)
            assert g.tuning == best
            assert best_ms <= default_ms * 1.05
            job.run(N.MODE_GRAPH)
            assert all(np.array_equal(a, b) for a, b in zip(job.download(), want))
        key = g.tuning_key
    assert key.startswith("gfx") and key.endswith("cu")
    chosen = dict(H.default_tuning(), rgb_rows=1, sao_parts=2, intra_per_cu=1)
    H.save_tuning(f, key, chosen)
    old = os.environ.get("HPVCD_TUNING_FILE")
    os.environ["HPVCD_TUNING_FILE"] = f
    try:
        with H.Gpu(0) as g2:
            assert g2.tuning == chosen
    finally:
        if old is None:
            del os.environ["HPVCD_TUNING_FILE"]
        else:
            os.environ["HPVCD_TUNING_FILE"] = old
    with H.Gpu(0) as g3:
        assert g3.tuning == H.default_tuning() or old is not None
