"""The allocation-free batched entry point (weno_batched_reconstruct_noalloc)
against the allocating one (weno_batched_reconstruct): same kernels, so the
outputs must be bitwise equal.  Two calls in a row on the shared scratch
array, the second with fewer and smaller blocks, for every direction, with
and without the fused limiter; then a third with the first set again."""
import ctypes as C

import numpy as np
import pytest

import wenostandalone_amd as wm
from tests.test_gpu_batched import Blocks

pytestmark = pytest.mark.gpu

LARGE = [(31, 22, 17, 5, 7), (15, 40, 12, 8, 5), (27, 9, 33, 2, 5), (12, 12, 12, 10, 7), (45, 14, 8, 3, 3),
         (19, 26, 21, 4, 7), (8, 8, 30, 6, 5)]
SMALL = [(9, 7, 11, 2, 5), (6, 10, 8, 3, 7)]


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield


def _both(specs, seed, d, flags):
    bl = Blocks(specs, seed=seed)
    bl.fill(2.5)
    wm.check(wm.lib.weno_batched_reconstruct(bl.blocks, len(specs), d, flags))
    ref = bl.download()
    bl.fill(2.5)
    wm.check(wm.lib.weno_batched_reconstruct_noalloc(bl.blocks, len(specs), d, flags))
    got = bl.download()
    for x in bl.keep:
        x.free()
    for fc in bl.faces:
        for L, R in fc:
            L.free()
            R.free()
    return ref, got


@pytest.mark.parametrize("flags", [wm.FUSE_COPY, wm.FUSE_COPY | wm.FUSE_MP, 0])
def test_noalloc_matches_batched(flags):
    for d in range(3):
        for specs, seed in ((LARGE, 10), (SMALL, 20), (LARGE, 30)):
            ref, got = _both(specs, seed, d, flags)
            for b in range(len(specs)):
                for k in range(2):
                    assert np.array_equal(ref[b][d][k], got[b][d][k], equal_nan=True), (flags, d, seed, b, k)


def test_noalloc_rejects_too_many_blocks():
    blocks = (wm.Block * 257)()
    rc = wm.lib.weno_batched_reconstruct_noalloc(blocks, 257, 0, wm.FUSE_COPY)
    assert rc == -1 and "at most 256" in wm._lib.last_error()
