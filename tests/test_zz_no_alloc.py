"""gvx_dist_pba_noalloc (the EDT in the library's fixed device workspace) against
gvx_dist_pba on the same inputs, bit-exact, through the C ABI.  Two calls in a
row share the workspace, the second on a smaller map; a map above the workspace
is rejected before anything runs."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GVX_ERR_BAD_ARG = 1


def _field(d, dens, seed):
    n = int(np.prod(np.array(d, dtype=np.int64)))
    idx = np.nonzero(np.random.default_rng(seed).random(n) < dens)[0]
    m = np.full(n, O.dist_uninit(d), np.uint32)
    x, y, z = idx % d[0], (idx // d[0]) % d[1], idx // (d[0] * d[1])
    m[idx] = (x | (y << 10) | (z << 20)).astype(np.uint32)
    return m


def _run(gvx, d, m, noalloc):
    g = gvx.Map(O.DIST_MAP, d, 0.01)
    g.write(m)
    if noalloc:
        gvx.check(gvx.lib().gvx_dist_pba_noalloc(g.h))
    else:
        gvx.check(gvx.lib().gvx_dist_pba(g.h))
    return g.read()


def test_noalloc_matches_pba_two_calls(gvx_lib):
    gvx = gvx_lib
    # 128^3 at 1 % (the workspace's stated size), then a smaller ragged map at 5 %:
    # the second call finds the first call's contents in the workspace
    for d, dens, seed in (((128, 128, 128), 0.01, 71), ((70, 33, 45), 0.05, 72)):
        m = _field(d, dens, seed)
        want = _run(gvx, d, m, noalloc=False)
        got = _run(gvx, d, m, noalloc=True)
        print("noalloc vs pba", d, dens, flush=True)
        assert np.array_equal(got, want), d
    # the small map checked against the brute-force oracle as well
    d = (21, 17, 13)
    m = _field(d, 0.03, 73)
    assert np.array_equal(_run(gvx, d, m, noalloc=True), O.edt_bruteforce(m, d))


def test_noalloc_rejects_maps_above_the_workspace(gvx_lib):
    gvx = gvx_lib
    d = (256, 256, 64)  # 4 Mi voxels: about 40 MiB of workspace
    g = gvx.Map(O.DIST_MAP, d, 0.01)
    with pytest.raises(gvx.GvxError) as e:
        gvx.check(gvx.lib().gvx_dist_pba_noalloc(g.h))
    assert e.value.code == GVX_ERR_BAD_ARG
    assert "workspace" in str(e.value)
