"""st_joined_to_held_static, the entry point that allocates nothing, against st_joined_to_held
(keep_gone = 1) through the C ABI: the same joined flags, bit for bit, and their count, on
two calls in a row, the second on a smaller structure (the scratch keeps the first call's
contents, so the entry point must clear what it reads); and a structure larger than its
scratch is refused with a message."""
import ctypes

import numpy as np
import pytest

import statics_hip as sh

pytestmark = pytest.mark.gpu


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def structure(ctx, n, seed):
    occ, theta = sh.generate(n, n, n, seed=seed, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx.upload(sites, held, e, g)
    # break about a third of the beams of a third of the cells, so that pieces come loose
    rng = np.random.default_rng(seed)
    cut = rng.random(len(sites)) < 0.3
    beams = np.where(cut, rng.integers(0, 0x8000, len(sites)) & 0x7FFF, 0x7FFF).astype(np.uint16)
    ctx.set_beams(beams)
    return len(sites)


def test_joined_static_matches_joined_to_held(ctx):
    L = sh.lib()
    L.st_joined_to_held_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    H = hip()
    for n, seed in ((48, 3), (24, 5)):
        structure(ctx, n, seed)
        want = ctx.joined_to_held(keep_gone=True)
        assert 0 < want.sum() < len(want)  # some nodes joined, some not
        nn = ctx.nodes
        dj, dc = ctypes.c_void_p(), ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(dj), nn) == 0 and H.hipMalloc(ctypes.byref(dc), 4) == 0
        junk = np.full(nn, 7, np.uint8)
        assert H.hipMemcpy(dj, junk.ctypes.data, nn, 1) == 0  # host to device
        rc = L.st_joined_to_held_static(ctx.h, None, dj, dc)
        assert rc == 0, ctx.L.st_last_error(ctx.h)
        got = np.zeros(nn, np.uint8)
        cnt = np.zeros(1, np.int32)
        assert H.hipDeviceSynchronize() == 0  # the call only enqueues on the context's stream
        assert H.hipMemcpy(got.ctypes.data, dj, nn, 2) == 0 and H.hipMemcpy(cnt.ctypes.data, dc, 4, 2) == 0
        H.hipFree(dj)
        H.hipFree(dc)
        assert (got == want).all(), (n, int((got != want).sum()))
        assert int(cnt[0]) == int(want.sum())
        # the context's gone set is untouched: st_joined_to_held with keep_gone agrees again
        assert (ctx.joined_to_held(keep_gone=True) == want).all()


def test_joined_static_refuses_beyond_its_scratch(ctx):
    L = sh.lib()
    L.st_joined_to_held_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    occ, theta = sh.generate(200, 200, 200, seed=2, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx.upload(sites, held, e, g)
    nn = ctx.nodes
    assert nn > (1 << 22)
    H = hip()
    dj = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dj), nn) == 0
    rc = L.st_joined_to_held_static(ctx.h, None, dj, None)
    assert H.hipDeviceSynchronize() == 0
    H.hipFree(dj)
    assert rc == sh.ST_BAD_INPUT
    assert b"scratch" in ctx.L.st_last_error(ctx.h)
