"""Two context pipelines (gr_config.pipelines = 2): consecutive asynchronous
draws alternate between two sets of buffers and streams. Frames must still
land in submission order, interleaved device fills and copies must stay
ordered with the draws, and every frame equals the oracle's."""
import json
import os

import numpy as np
import pytest

from gorender_amd import scenes as S
from parity import digest
from scene_cases import config4_batch

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_digests.json")))


def read(buf, fr):
    return buf.to_numpy(np.uint8, (fr.height, fr.width, 4))


def test_same_buffers_keep_submission_order(ctx):
    """Eight draws of two different scenes into one pair of buffers, with a fill
    between some of them: the buffers end with the last draw."""
    assert ctx.config().pipelines == 2
    a = S.config1("ref")
    b = S.config1("cfg")
    ha = ctx.upload_scene(a)[0]
    hb = ctx.upload_scene(b)[0]
    fr = a.frame
    rg, dp = ctx.frame_buffers(fr.width, fr.height)
    want = {}
    for name, sc, h in (("a", a, ha), ("b", b, hb)):
        ctx.draw_device(sc, h, rg.ptr, dp.ptr)
        ctx.synchronize()
        want[name] = read(rg, fr).copy()
    assert not np.array_equal(want["a"], want["b"])
    for last in ("a", "b"):
        seq = ["a", "b", "a", "b", "a", "b", "a", "b"]
        if last == "a":
            seq = seq[1:] + ["a"]
        for k, name in enumerate(seq):
            if k % 3 == 2:
                rg.fill(0x5A)  # ordered after the previous draw, before the next
            ctx.draw_device(a if name == "a" else b, ha if name == "a" else hb, rg.ptr, dp.ptr)
        ctx.synchronize()
        assert np.array_equal(read(rg, fr), want[last]), f"buffers do not hold draw {last}"
    # A fill after the last draw wins over it.
    ctx.draw_device(a, ha, rg.ptr, dp.ptr)
    rg.fill(0x11)
    ctx.synchronize()
    assert (read(rg, fr) == 0x11).all()
    rg.free()
    dp.free()


def test_back_to_back_batches_match_oracle(ctx):
    """Config 4's 16 frames as four batches of four, submitted without waiting
    (two pipes in flight), each frame into its own buffers."""
    frames = config4_batch()
    handles = [ctx.upload_scene(s)[0] for s in frames]
    outs = [ctx.frame_buffers(s.frame.width, s.frame.height) for s in frames]
    for _ in range(2):  # the first round may grow the tile-list capacity
        for q in range(4):
            sl = slice(4 * q, 4 * q + 4)
            ctx.draw_batch(frames[sl], handles[sl], [(o[0].ptr, o[1].ptr) for o in outs[sl]])
        try:
            ctx.synchronize()
            break
        except Exception:
            continue
    for k, s in enumerate(frames):
        h, w = s.frame.height, s.frame.width
        rgba = outs[k][0].to_numpy(np.uint8, (h, w, 4))
        depth = outs[k][1].to_numpy(np.float32, (h, w))
        got = digest(rgba, depth, {"tpf": 0, "triangles": 0, "fragments_covered": 0,
                                   "faces_front": 0, "pixels_shaded": 0})
        want = GOLDEN[f"config4-{k}"]
        assert (got["rgba"], got["depth"]) == (want["rgba"], want["depth"]), f"config4-{k} differs"
        outs[k][0].free()
        outs[k][1].free()
