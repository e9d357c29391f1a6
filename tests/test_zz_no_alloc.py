"""gu_lbs_noalloc, the allocation-free form of gu_lbs (its scratch is a
fixed device array per GPU), against gu_lbs through the C ABI.

Two calls in a row on the same GPU, the second with a smaller input, on the
one-launch path (up to 8192 segments) and on the three-phase path: the
second call must not see the first call's leftovers in the shared scratch.
Outputs must equal gu_lbs's bit for bit (and the oracle's).  Inputs beyond
the scratch's limits are refused with GU_ERR_CAPACITY.
"""
import ctypes

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

SENTINEL = -7


def _counts(S, hi, seed):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, hi, S).astype(np.int32)
    c[rng.random(S) < 0.1] = 0
    return c


def _run(gu, fn, c):
    S = len(c)
    W = int(c.astype(np.int64).sum())
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer.from_array(np.full(S, SENTINEL, np.int32))
    rb = gu.DeviceBuffer.from_array(np.full(W + 64, SENTINEL, np.int32))  # 64 guard entries
    try:
        Wd, _ = fn(cb, S, sb, rb, W)
        return Wd, sb.to_host(np.int32, S), rb.to_host(np.int32, W + 64)
    finally:
        for b in (cb, sb, rb):
            b.close()


@pytest.mark.parametrize("sizes", [(8000, 3000), (300000, 100000), (300000, 5000), (6000, 200000)])
def test_noalloc_matches_gu_lbs_twice(gu, orc, gpu, sizes):
    for k, S in enumerate(sizes):
        c = _counts(S, 200 if S > 8192 else 4000, seed=S + k)
        announce("gu_lbs_noalloc call %d of %s: S=%d W=%d" % (k + 1, sizes, S, int(c.sum())))
        W1, scan1, res1 = _run(gu, gu.lbs_device_noalloc, c)
        W0, scan0, res0 = _run(gu, gu.lbs_device, c)
        assert W1 == W0 == int(c.sum())
        assert np.array_equal(scan1, scan0)
        assert np.array_equal(res1, res0)  # guard entries included: nothing written past W
        assert np.all(res1[W1:] == SENTINEL)
        oscan, oW = orc.exclusive_scan(c)
        assert np.array_equal(scan1, oscan) and np.array_equal(res1[:W1], orc.lbs(oW, oscan))


def test_noalloc_scan_only_and_limits(gu, gpu):
    c = _counts(20000, 100, seed=3)
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(4 * len(c))
    announce("gu_lbs_noalloc scan only, then the size limits")
    try:
        W, _ = gu.lbs_device_noalloc(cb, len(c), sb, None, 0)
        assert W == int(c.sum())
        assert np.array_equal(sb.to_host(np.int32, len(c)),
                              np.concatenate([[0], np.cumsum(c)[:-1]]).astype(np.int32))
        # more segments than the scratch holds: refused before any device work
        with pytest.raises(gu.GuError) as e:
            gu.lbs_device_noalloc(cb, gu.LBS_NOALLOC_MAX_SEGMENTS + 1, sb, None, 0)
        assert e.value.status == gu.ERR_CAPACITY
    finally:
        cb.close()
        sb.close()
    # more items than its tile-start table: the scan and W are written, no
    # result entry is
    big = np.array([gu.LBS_NOALLOC_MAX_ITEMS, 1], dtype=np.int32)
    W = int(big.astype(np.int64).sum())
    cb = gu.DeviceBuffer.from_array(big)
    sb = gu.DeviceBuffer(8)
    rb = gu.DeviceBuffer(4 * W)
    try:
        rb.zero()
        Wo = ctypes.c_int64()
        cfg, st = gu.make_config(), gu.Stats()
        rc = gu.lib().gu_lbs_noalloc(cb.ptr, 2, sb.ptr, rb.ptr, W, ctypes.byref(Wo), ctypes.byref(cfg),
                                     ctypes.byref(st))
        assert rc == gu.ERR_CAPACITY and Wo.value == W
        assert np.array_equal(sb.to_host(np.int32, 2), [0, gu.LBS_NOALLOC_MAX_ITEMS])
        assert not rb.to_host(np.int32).any()
    finally:
        for b in (cb, sb, rb):
            b.close()
