"""Host-side argument checks of the batch entry point (no GPU needed: every
case is rejected before anything is enqueued)."""

import ctypes as C

from mkpp_amd import native


def _seg(name=b"saprc99", cells=10):
    # device pointers are never dereferenced on these paths; any non-null value will do
    return native.BatchSegment(name, 0x1000, cells, cells, None, None, None, 0, 0x2000, 0x3000, None, None, None)


def _call(segs, n=None, dt=3600.0, steps=1, method=3, seed=1e-6, cpl=1, wg=0, ws=0x4000):
    arr = (native.BatchSegment * max(len(segs), 1))(*segs)
    return native.lib().mkpp_amd_device_integrate_batch(arr if segs else None, len(segs) if n is None else n, dt,
                                                        steps, method, seed, 100000, cpl, wg, C.c_void_p(ws), None)


def test_rejects_bad_arguments():
    assert _call([]) == -101  # null segment array
    assert _call([_seg()], ws=0) == -101
    assert _call([_seg()], n=0) == -103
    assert _call([_seg()] * 9) == -103  # more than MKPP_AMD_BATCH_MAX_SEGMENTS
    assert _call([_seg()], dt=0.0) == -103
    assert _call([_seg()], dt=float("nan")) == -103
    assert _call([_seg()], steps=0) == -103
    assert _call([_seg()], method=7) == -103
    assert _call([_seg()], cpl=0) == -103
    assert _call([_seg()], wg=-1) == -103
    assert _call([_seg(b"nosuch")]) == -106
    bad = _seg()
    bad.stride = 5  # stride < num_cells
    assert _call([bad]) == -103
    nul = _seg()
    nul.atol = None
    assert _call([nul]) == -101


def test_batch_engines_cover_the_batched_mechanisms():
    from mkpp_amd.gen.model import BATCHED

    names = native.mechanisms()
    assert set(BATCHED) <= set(names)
    for nm in names:
        # a batched mechanism passes the name check and fails later only on the
        # bad stride; a registered mechanism outside the batch kernel is not found
        s = _seg(nm.encode())
        s.stride = 1
        assert _call([s]) == (-103 if nm in BATCHED else -106)
