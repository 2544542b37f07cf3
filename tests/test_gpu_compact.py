"""The compact schedule (cap and compact, vrt_cast_ex with VRT_SCHED_COMPACT): its per-stream survivor
state under graph capture, growth and reuse, concurrent streams, and the device ray count."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal, blob_world, random_rays

pytestmark = pytest.mark.gpu

COMPACT = 3


@pytest.fixture(scope="module")
def blob_pair(vrt, oracle_mod):
    from test_oracle import dense_world

    vox = blob_world(128, seed=61, density=0.008)
    return vrt.World.from_scene(vrt.Scene.from_dense(vox)), dense_world(oracle_mod, vox)


def rays_and_ref(ow, n, seed, max_iters=128, W=(0, 0, 0)):
    from voxelrt_amd import RayBatch

    o, d = random_rays(n, 128, seed=seed, inside=False)
    o = (o - np.array(W, np.float32)[:, None]).astype(np.float32)
    return RayBatch.from_numpy(o, d, W), ow.cast(o, d, W, max_iters, with_pos=True)


def check(h, ref, what):
    torch.cuda.synchronize()
    assert_hits_equal(h.t.cpu().numpy(), h.rec.cpu().numpy(), ref["t"], ref["rec"], what)
    assert np.array_equal(h.t.cpu().numpy().view(np.uint32), ref["t"].view(np.uint32)), what
    hit = (ref["rec"][:, 3] & 0xFF) != 0
    assert np.array_equal(h.pos.cpu().numpy()[:, hit].view(np.uint32), ref["pos"][:, hit].view(np.uint32)), what


def test_graph_replays_match_eager(vrt, cuda, blob_pair):
    """A compact cast captured in a graph (after one eager cast on the capture stream has sized
    its survivor state), replayed three times with other compact casts on the same stream in
    between: every replay writes the oracle's hits."""
    import voxelrt_amd as V

    world, ow = blob_pair
    rb, ref = rays_and_ref(ow, 30000, seed=62)
    rb2, ref2 = rays_and_ref(ow, 20000, seed=63)
    s = torch.cuda.Stream()
    h = V.HitBatch.empty(rb.n, with_pos=True)
    h2 = V.HitBatch.empty(rb2.n, with_pos=True)
    with torch.cuda.stream(s):
        world.cast(rb, 128, with_pos=True, schedule=COMPACT, out=h, stream=s)
    torch.cuda.synchronize()
    check(h, ref, "eager")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        world.cast(rb, 128, with_pos=True, schedule=COMPACT, out=h, stream=s)
    for rep in range(3):
        h.rec.fill_(-1)
        h.t.fill_(0.0)
        h2.rec.fill_(-1)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            world.cast(rb2, 128, with_pos=True, schedule=COMPACT, pass_cap=5 + rep, out=h2, stream=s)
        torch.cuda.synchronize()
        check(h2, ref2, f"eager between replays {rep}")
        g.replay()
        torch.cuda.synchronize()
        check(h, ref, f"replay {rep}")


def test_survivor_state_grows_and_is_reused(vrt, cuda, blob_pair):
    """Casts of 1 000, 60 000 (the state grows) and 1 000 rays again on one stream, at each
    world-origin specialization."""
    import voxelrt_amd as V

    world, ow = blob_pair
    s = torch.cuda.Stream()
    for k, (n, W) in enumerate([(1000, (0, 0, 0)), (60000, (-7, 3, 11)), (1000, (32, -64, 96))]):
        rb, ref = rays_and_ref(ow, n, seed=70 + k, W=W)
        with torch.cuda.stream(s):
            h = world.cast(rb, 128, with_pos=True, schedule=COMPACT, pass_cap=16, stream=s)
        check(h, ref, f"n={n} W={W}")


def test_two_streams_at_once(vrt, cuda, blob_pair):
    """Compact casts queued on two streams without a synchronization between them use their own
    survivor state."""
    world, ow = blob_pair
    rb_a, ref_a = rays_and_ref(ow, 40000, seed=81)
    rb_b, ref_b = rays_and_ref(ow, 40000, seed=82)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(2):  # the first round sizes both streams' state
        with torch.cuda.stream(sa):
            ha = world.cast(rb_a, 128, with_pos=True, schedule=COMPACT, stream=sa)
        with torch.cuda.stream(sb):
            hb = world.cast(rb_b, 128, with_pos=True, schedule=COMPACT, stream=sb)
        check(ha, ref_a, "stream a")
        check(hb, ref_b, "stream b")


def test_capture_before_the_state_is_sized_is_refused(vrt, cuda, blob_pair):
    """Under capture the schedule cannot allocate: a stream whose survivor state is smaller than the ray
    count gets VRT_ERR_STATE and launches nothing."""
    import voxelrt_amd as V

    world, ow = blob_pair
    rb, _ = rays_and_ref(ow, 5000, seed=91)
    s = torch.cuda.Stream()
    h = V.HitBatch.empty(rb.n, with_pos=True)
    g = torch.cuda.CUDAGraph()
    with pytest.raises(V.VrtError, match="before capturing"):
        with torch.cuda.graph(g, stream=s):
            world.cast(rb, 128, with_pos=True, schedule=COMPACT, out=h, stream=s)
    torch.cuda.synchronize()


@pytest.mark.parametrize("k", [0, 1, 777, 30000])
def test_device_count(vrt, cuda, blob_pair, k):
    """With a device ray count the first pass traces exactly that many rays; the rest of the
    output stays untouched and the counters count k rays."""
    import voxelrt_amd as V

    world, ow = blob_pair
    rb, ref = rays_and_ref(ow, 30000, seed=95)
    cnt = torch.tensor([k], dtype=torch.int32, device="cuda")
    h = V.HitBatch.empty(rb.n)
    h.rec.fill_(-7)
    ctr = V.new_counters()
    world.cast(rb, 128, schedule=COMPACT, pass_cap=8, count=cnt, out=h, counters=ctr)
    torch.cuda.synchronize()
    rec = h.rec.cpu().numpy()
    assert int(ctr[0].item()) == k
    assert np.array_equal(rec[:k], ref["rec"][:k])
    assert np.all(rec[k:] == -7)


def test_release_stream_frees_and_later_casts_reallocate(vrt, cuda, blob_pair):
    """vrt_stream_release frees a stream's compact and persistent state; casts on the stream
    afterwards allocate it again and write the oracle's hits; releasing a stream without
    state, or twice, is a no-op."""
    import voxelrt_amd as V

    world, ow = blob_pair
    rb, ref = rays_and_ref(ow, 20000, seed=97)
    s = torch.cuda.Stream()
    for _ in range(2):
        with torch.cuda.stream(s):
            h1 = world.cast(rb, 128, with_pos=True, schedule=COMPACT, pass_cap=12, stream=s)
            h2 = world.cast(rb, 128, with_pos=True, schedule=2, chunk=256, stream=s)
        V.release_stream(s)
        check(h1, ref, "compact before release")
        check(h2, ref, "persistent before release")
    V.release_stream(s)
    V.release_stream(torch.cuda.Stream())
