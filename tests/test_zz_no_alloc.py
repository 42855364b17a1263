"""eolc_collide_noalloc (no device allocation; its buffers come from the library's
per-GPU scratch array) against eolc_collide on the same inputs, through the C ABI:
record sets equal bit for bit, on two calls in a row with the second smaller (a
smaller cloth, fewer obstacles, CD instead of CD2), with an allocating call between
them; and a mesh whose query cannot fit the scratch is refused with a message
before any device work."""
import numpy as np
import pytest

import eolc
from eolc import scene
from test_gpu_collide import rec_tuple

pytestmark = pytest.mark.gpu


def same_records(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert rec_tuple(x) == rec_tuple(y), f"record {k} differs"


def test_noalloc_matches_collide_twice_second_smaller():
    with eolc.Context(0) as c:
        calls = []
        for n, P, B, eol in ((128, 64, 4, False), (48, 16, 2, True)):
            sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
            pts, nrm, whd, E = scene.obstacles(seed=1236 + n, P=P, B=B)
            # boxes low enough to reach the sheet; points within the query band
            c.set_topology(sc.X, sc.tri, sc.hinge)
            c.upload_state(sc.x0, sc.v0)
            print(f"noalloc vs collide: {n}x{n}, {P} points, {B} boxes, eol {eol}", flush=True)
            got = c.collide(1e-3, pts, nrm, whd, E, eol_points=eol, noalloc=True)
            ref = c.collide(1e-3, pts, nrm, whd, E, eol_points=eol)
            again = c.collide(1e-3, pts, nrm, whd, E, eol_points=eol, noalloc=True)
            kinds = sorted({(r.count1, r.count2) for r in ref})
            print(f"  {len(ref)} records, kinds {kinds}", flush=True)
            assert len(ref) > 0
            same_records(got, ref)
            same_records(again, ref)
            calls.append(len(ref))
        assert calls[1] < calls[0]


def test_noalloc_refuses_a_mesh_beyond_the_scratch():
    sc = scene.Scene(1024)
    with eolc.Context(0) as c:
        c.set_topology(sc.X, sc.tri, sc.hinge)
        c.upload_state(sc.x0, sc.v0)
        with pytest.raises(eolc.EolcError) as ei:
            c.collide(1e-3, noalloc=True)
        assert ei.value.code == eolc.ERR_INVALID and "scratch" in str(ei.value)
        # the context still works after the refusal
        small = scene.Scene(16, sigma_x=1e-3)
        c.set_topology(small.X, small.tri, small.hinge)
        c.upload_state(small.x0, small.v0)
        pts, nrm, whd, E = scene.obstacles(seed=9, P=8, B=1)
        same_records(c.collide(1e-3, pts, nrm, whd, E, noalloc=True), c.collide(1e-3, pts, nrm, whd, E))
