"""The serving loop on the GPU (hpvcd_server_*).

Requests of mixed sizes and formats arrive from several threads at once and
are continuously batched: every image must equal the image the same file
gives decoded alone (hpvcd_decode_heic), bit for bit, whatever batch it
rode in; batches must actually mix requests; and a batch cap of one picture
(admission refusing every second request) must give the same images.
"""
import threading

import numpy as np
import pytest

import hpvcd_amd as H
from hpvcd_amd import N

pytestmark = pytest.mark.gpu

MIX = [
    dict(H.WORKLOADS["single512"], tile_width=256, tile_height=256, out_width=256, out_height=256, base_seed=301),
    dict(H.WORKLOADS["grid4032"], grid_cols=2, grid_rows=2, out_width=1024, out_height=768, base_seed=303),
    dict(H.WORKLOADS["grid4032"], grid_cols=4, grid_rows=4, out_width=2048, out_height=2048, bit_depth=10,
         base_seed=305),
    dict(H.WORKLOADS["rext4096"], grid_cols=2, grid_rows=1, out_width=1024, out_height=512, base_seed=307),
    dict(H.WORKLOADS["single512"], chroma_idc=0, base_seed=309),
    dict(H.WORKLOADS["grid4032"], grid_cols=2, grid_rows=1, out_width=1000, out_height=500,
         orientation=N.ORIENT_ROTATE_90, base_seed=311),
]


@pytest.fixture(scope="module")
def files():
    fs = [H.gen_heic(**p) for p in MIX]
    with H.Decoder() as d:
        alone = [d.decode(f) for f in fs]
        alone8 = [d.decode_rgb8(f) for f in fs]
    return fs, alone, alone8


def _run(server, fs, order, rgb8=False, threads=4):
    results = {}
    lock = threading.Lock()

    def worker(idx):
        for i in idx:
            t = server.submit(fs[order[i]], rgb8=rgb8)
            img, st = server.wait(t)
            with lock:
                results[i] = (order[i], img, st)

    chunks = [list(range(k, len(order), threads)) for k in range(threads)]
    ths = [threading.Thread(target=worker, args=(c,)) for c in chunks]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return results


def test_mixed_requests_equal_alone(files):
    fs, alone, alone8 = files
    order = [i % len(fs) for i in range(24)]
    print("serving: 24 requests from 4 threads", flush=True)
    with H.Server() as srv:
        res = _run(srv, fs, order)
        res8 = _run(srv, fs, order[:12], rgb8=True)
        _, reqs, _ = srv.totals()
    for i, (k, img, st) in res.items():
        assert np.array_equal(img, alone[k]), (i, k)
        assert st["latency_ms"] > 0 and st["batch_requests"] >= 1
    for i, (k, img, st) in res8.items():
        assert np.array_equal(img, alone8[k]), (i, k)
    assert reqs == 36


def test_requests_parsed_together_share_launches(files):
    """A burst of requests with a batching delay: a free slot waits for the
    requests still in the parse, so launches hold several requests; every
    image still equals its file decoded alone."""
    fs, alone, _ = files
    order = [i % len(fs) for i in range(18)]
    print("serving: burst of 18 with a 300 ms batching delay", flush=True)
    with H.Server(batch_delay_us=300_000) as srv:
        tickets = [srv.submit(fs[k]) for k in order]
        out = [srv.wait(t) for t in tickets]
        batches, reqs, _ = srv.totals()
    for k, (img, st) in zip(order, out):
        assert np.array_equal(img, alone[k])
    assert reqs == 18 and batches < reqs
    assert max(st["batch_requests"] for _, st in out) > 1


def test_batch_cap_one_picture_same_images(files):
    fs, alone, _ = files
    order = [0, 4, 0, 4, 1]
    print("serving: one-picture batches", flush=True)
    with H.Server(max_batch_pictures=1) as srv:
        res = _run(srv, fs, order, threads=2)
        batches, reqs, _ = srv.totals()
    for i, (k, img, st) in res.items():
        assert np.array_equal(img, alone[k])
    # single-picture requests never share a batch under the cap
    assert batches == reqs


def test_bad_request_fails_alone(files):
    fs, alone, _ = files
    with H.Server() as srv:
        bad = srv.submit(b"\x00\x00\x00\x0cftypmp42")
        good = srv.submit(fs[0])
        with pytest.raises(H.HpvcdError) as e:
            srv.wait(bad)
        assert e.value.status == N.E_NOT_HEIF
        img, _ = srv.wait(good)
        assert np.array_equal(img, alone[0])
