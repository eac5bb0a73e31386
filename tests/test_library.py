"""The C ABI without a GPU: the library loads, exports what include/statics_hip.h
declares, and its host-only parts (the generator) work."""
import ctypes
import os
import re

import numpy as np
import pytest

import statics_hip as sh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "statics_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(st_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(sh.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(sh.EXPORTS)


def test_exports_every_mechanics_symbol():
    """include/mechanics_hip.h's entry points are exported and bound."""
    with open(os.path.join(ROOT, "include", "mechanics_hip.h")) as f:
        names = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(me_\w+)\s*\(", f.read(), re.M)))
    lib = ctypes.CDLL(sh.LIB_PATH)
    assert len(names) == 6
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(sh.ME_EXPORTS)


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(sh.StaticsError):
        sh.Context()


def test_generator_is_seeded_and_shaped():
    a, ta = sh.generate(32, 32, 32, seed=5, sigma=0.5)
    b, tb = sh.generate(32, 32, 32, seed=5, sigma=0.5)
    c, _ = sh.generate(32, 32, 32, seed=6, sigma=0.5)
    assert (a == b).all() and (ta == tb).all() and not (a == c).all()
    assert (a[:, :, 0] == 2).all() and not (a[:, :, 1:] == 2).any()  # held ground layer only
    fill = (a > 0).mean()
    assert 0.15 < fill < 0.45
    # pillars run from z = 1 to the top: some column is full above the ground
    assert (a[:, :, 1:] == 1).all(axis=2).any()
    assert abs(ta.std() - 0.5) < 0.02 and abs(ta.mean()) < 0.02


def test_generator_rejects_bad_params():
    p = sh.StGenParams(0, 4, 4, 0, 8.0, 2, 6, 8, 16, 0.2, 0.0)
    occ = np.zeros(1, np.uint8)
    assert sh.lib().st_generate(ctypes.byref(p), sh._p(occ), None) == sh.ST_BAD_INPUT


def test_struct_layouts_match_the_header():
    """The binding's structures have the sizes and offsets the header states (and the
    library asserts at compile time)."""
    assert ctypes.sizeof(sh.StSolveOpts) == 32
    assert [getattr(sh.StSolveOpts, f).offset for f in
            ("precision", "tol", "max_iter", "refresh", "warm_start", "release_hierarchy")] == [0, 8, 16, 20, 24, 28]
    assert ctypes.sizeof(sh.StStats) == 88
    assert [getattr(sh.StStats, f).offset for f in
            ("iters", "levels", "refreshes", "status", "rel_residual", "setup_ms", "solve_ms", "nodes", "dof_nodes",
             "cells", "bricks", "retries", "first_status", "mem_peak")] == \
        [0, 4, 8, 12, 16, 24, 32, 40, 48, 56, 64, 72, 76, 80]
    with open(os.path.join(ROOT, "include", "statics_hip.h")) as f:
        text = f.read()
    assert "StSolveOpts is 32 bytes" in text and "StStats is 88 bytes" in text
