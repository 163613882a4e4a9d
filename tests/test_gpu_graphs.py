"""Captured bounce-loop graphs, launch-shape settings and the autotuner.

None of them may change a bit of the image: each case compares the binary32
running mean (and the binary64 sums where enabled) bit for bit against plain
launches or the CPU oracle.  Graph replays are checked one by one, with host
reads (eager copies and synchronisations) between them.
"""
import numpy as np
import pytest

import oracle
from bluetruthengine_amd import renderer, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _renders(sd, calls, frames, **cfg):
    r = renderer.Renderer(sd, **cfg)
    r.enable_accum64(True)
    out, launches = [], []
    for _ in range(calls):
        r.render(frames)
        launches.append(r.stats().graphLaunches)
        out.append((r.accum().copy(), r.accum64().copy()))
    return out, launches


def test_graph_replays_match_plain_launches():
    _say("case: Cornell 64x64 depth 8, 5 calls x 2 frames, graphs vs plain launches")
    sd = scenes.cornell_c1(width=64, height=64, spp=10, max_depth=8)
    plain, l0 = _renders(sd, 5, 2, graphs=0)
    graph, l1 = _renders(sd, 5, 2, graphs=1)
    assert all(n == 0 for n in l0)
    assert all(n >= 1 for n in l1), l1
    for i, ((a, a64), (b, b64)) in enumerate(zip(plain, graph)):
        assert np.array_equal(a, b), f"call {i}: running mean differs"
        assert np.array_equal(a64, b64), f"call {i}: binary64 sums differ"


def test_graph_replays_match_oracle():
    _say("case: Cornell 48x48 depth 8, 3 calls x 2 frames with graphs, against the oracle")
    sd = scenes.cornell_c1(width=48, height=48, spp=6, max_depth=8)
    graph, _ = _renders(sd, 3, 2, graphs=1)
    ref = oracle.OracleScene(sd)
    racc = ref.render(48, 48, 0, 6, sd.max_depth)
    assert np.array_equal(graph[-1][0], racc)


def test_deep_paths_run_in_segments_with_early_exit():
    # depth 40: three captured segments of at most 16 bounces; the host stops
    # once a segment leaves no path, which Russian roulette guarantees here
    _say("case: Cornell 48x48 depth 40, 2 calls x 2 frames, graphs vs plain vs oracle")
    sd = scenes.cornell_c1(width=48, height=48, spp=4, max_depth=40)
    plain, _ = _renders(sd, 2, 2, graphs=0)
    graph, _ = _renders(sd, 2, 2, graphs=1)
    for (a, a64), (b, b64) in zip(plain, graph):
        assert np.array_equal(a, b) and np.array_equal(a64, b64)
    racc = oracle.OracleScene(sd).render(48, 48, 0, 4, sd.max_depth)
    assert np.array_equal(graph[-1][0], racc)


@pytest.mark.parametrize("cfg", [dict(traversal_stack=8), dict(traversal_stack=32), dict(work_chunk=1),
                                 dict(work_chunk=8), dict(traversal_blocks_per_cu=2), dict(bvh_width=2),
                                 dict(bvh_width=2, traversal_stack=8, work_chunk=1),
                                 dict(bvh_width=8), dict(bvh_width=8, traversal_stack=8, work_chunk=2)])
def test_launch_shape_does_not_change_result(cfg):
    _say(f"case: reduced env-lit 128x72 depth 12, 2 frames, launch shape {cfg} vs default")
    sd = scenes.small_envlit()
    base, _ = _renders(sd, 1, 2)
    other, _ = _renders(sd, 1, 2, **cfg)
    assert np.array_equal(base[0][0], other[0][0])
    assert np.array_equal(base[0][1], other[0][1])


def test_autotune_keeps_result_and_valid_config():
    _say("case: autotune on Cornell 128x128 depth 8, 2-frame passes")
    sd = scenes.cornell_c1(width=128, height=128, spp=2, max_depth=8)
    r = renderer.Renderer(sd)
    ms, cfg = r.autotune(2)
    _say(f"autotune: best {ms:.3f} ms with {cfg}")
    assert ms > 0.0
    assert cfg["traversalStack"] in (8, 16, 32)
    assert cfg["workChunk"] in (1, 2, 4, 8)
    assert 0 <= cfg["traversalBlocksPerCU"] <= 64
    assert cfg["bvhWidth"] in (2, 4, 8)
    assert r.frame_index() == 0
    r.render(2)
    ref = oracle.OracleScene(sd).render(128, 128, 0, 2, sd.max_depth)
    assert np.array_equal(r.accum(), ref)
