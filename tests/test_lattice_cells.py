"""st_lattice_cells (host only, no GPU): a lattice as the reference holds it back to the cells
lattice::build made it from, against the cells the oracle built it from; refusal of lattices
the device's operator cannot take (gathered ones, broken beam geometry); per-beam strengths
of bodies::as_generated lattices."""
import numpy as np
import pytest

import oracle as orc
import statics_hip as sh
from harness import FACES, PAIR_OF, PAIRS, mats_from


def built(n=10, seed=3, level=0):
    occ, theta = sh.generate(n, n, n, seed=seed, sigma=0.5, slab_min=4, slab_max=6)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    mats = mats_from(e, g, rho=1.5, strength=(2.0, 3.0, 4.0))
    return orc.Lattice(level, sites.astype(np.int64) << level, mats, held), sites, held, mats


@pytest.mark.parametrize("level", [0, 2])
def test_cells_of_a_built_lattice(level):
    L, sites, held, mats = built(level=level)
    c = sh.lattice_cells(L)
    assert c["spacing"] == orc.UNIT * (1 << level)
    got = {tuple(s): k for k, s in enumerate(c["sites"].tolist())}
    # every free cell is found, with its stiffness and material
    for i in np.flatnonzero(held == 0):
        k = got[tuple(sites[i].tolist())]
        assert not c["held"][k] and c["e"][k] == mats[i, 1] and c["g"][k] == mats[i, 2] and c["rho"][k] == 1.5
        assert (c["strength"][k][c["beams"][k] & (1 << np.arange(15)) != 0] == [2.0, 3.0, 4.0]).all()
    # the held cells found are exactly the held ones next to a free cell
    free = {tuple(s) for s, h in zip(sites.tolist(), held) if not h}
    want_held = {tuple(s) for s, h in zip(sites.tolist(), held)
                 if h and any((s[0] + d[0], s[1] + d[1], s[2] + d[2]) in free for d in FACES.tolist())}
    assert {tuple(s) for s, h in zip(c["sites"].tolist(), c["held"]) if h} == want_held
    # each bond's beam: its cell (the oracle's beam_cell) and its pair
    cell_of_site = {tuple(s): i for i, s in enumerate(sites.tolist())}
    bi = c["bond_index"]
    assert (bi >= 0).all() and (bi % 4 == np.tile(np.arange(4), L.n_bonds // 4)).all()  # build's order
    bob = bi // 4
    cell_dev = bob // 15
    assert (bob[0::4] == bob[1::4]).all() and (bob[0::4] == bob[3::4]).all()
    for b in range(0, L.n_bonds, 4 * 7):
        i_or = L.beam_cell[b // 4]
        assert cell_of_site[tuple(c["sites"][cell_dev[b]].tolist())] == i_or
    # held-held beams are absent, the others present
    nbeams = np.unpackbits(c["beams"].view(np.uint8)).sum()
    assert nbeams == L.n_bonds // 4


def test_gathered_lattice_needs_sections():
    """Without room for the sections, a gathered lattice is refused (its beams have sections
    of their own, statics/src/gather.rs:386-390); a built one is not."""
    L, *_ = built(n=12)
    G = L.gathered()
    with pytest.raises(sh.StaticsError) as e:
        sh.lattice_cells(G, sections=False)
    assert e.value.code == sh.ST_BAD_INPUT and "section" in str(e.value)
    c = sh.lattice_cells(L, sections=False)
    assert "section" not in c and c["beams"].any()


@pytest.mark.parametrize("levels", [1, 2])
def test_gathered_lattice_cells(levels):
    """A lattice gathered once and twice: each beam found in its cell at the pair its end nodes
    are, with its own section (4 a over SECTION s^2) and strengths, and its cell's e and g."""
    L, *_ = built(n=16, seed=4)
    for _ in range(levels):
        L = L.gathered()
    c = sh.lattice_cells(L)
    s = sh.UNIT * 2 ** L.level
    assert c["spacing"] == s
    bi = c["bond_index"]
    flip = bi < 0
    at = np.where(flip, -1 - bi, bi)
    cell, q = at // 60, at // 4 % 15
    # a beam's four bonds: one beam, all its bonds one way, the places 0..3 each once
    assert (cell.reshape(-1, 4) == cell[0::4, None]).all() and (q.reshape(-1, 4) == q[0::4, None]).all()
    assert (flip.reshape(-1, 4) == flip[0::4, None]).all() and flip.any() and not flip.all()
    assert (np.sort((at % 4).reshape(-1, 4), axis=1) == np.arange(4)).all()
    # run from the greater face, the first two bonds swap places (n and e1 negated, e2 kept)
    want = np.tile(np.arange(4), L.n_bonds // 4)
    want = np.where(flip & (want < 2), want ^ 1, want)
    assert (at % 4 == want).all()
    assert (c["beams"][cell] >> q & 1).all()
    assert sum(bin(int(b)).count("1") for b in c["beams"]) == L.n_bonds // 4
    sec = c["section"][cell, q]
    np.testing.assert_allclose(sec * 0.336 * s * s, 4 * L.a, rtol=1e-15)
    assert np.ptp(sec) > 0.5  # the sections differ beam by beam
    np.testing.assert_array_equal(c["strength"][cell, q], L.bond_mat[:, 3:6])
    np.testing.assert_array_equal(c["e"][cell], L.bond_mat[:, 1])
    np.testing.assert_array_equal(c["g"][cell], L.bond_mat[:, 2])
    # beams' ends: the two faces of the pair at the cell's site
    site = c["sites"][cell].astype(np.int64)
    for end in (0, 1):
        v = np.where(flip, L.ij[:, 1 - end], L.ij[:, end])  # the node at the pair's lesser face, then greater
        f = np.array([FACES[PAIRS[int(t)][end]] for t in q])
        lo = site + np.minimum(f, 0)
        hi = lo + np.abs(f)
        np.testing.assert_array_equal(L.around[v, 0] >> L.level, lo)
        np.testing.assert_array_equal(L.around[v, 1] >> L.level, hi)


def test_broken_beams_are_refused():
    L, *_ = built()
    arr, _ = sh.lattice_arrays(L)

    class Lat:
        pass

    def with_(**kw):
        lat = Lat()
        lat.level = L.level
        for k, v in arr.items():
            setattr(lat, k, kw.get(k, v))
        return lat

    ij = arr["ij"].copy()
    ij[1], ij[4] = ij[4].copy(), ij[1].copy()  # a bond of the second beam in the first
    with pytest.raises(sh.StaticsError, match="one beam's four"):
        sh.lattice_cells(with_(ij=ij))
    p = arr["p"].copy()
    p[[0, 1]] = p[[1, 0]]  # +r e1 and -r e1 swapped: bond forces would come out in another order
    with pytest.raises(sh.StaticsError, match="in order"):
        sh.lattice_cells(with_(p=p))
    m = arr["bond_mat"].copy()
    m[0:4, 1] *= 2.0  # one beam of a cell stiffer than the others
    with pytest.raises(sh.StaticsError, match="stiffness"):
        sh.lattice_cells(with_(bond_mat=m))
    m = arr["bond_mat"].copy()
    m[2, 3] *= 2.0  # one bond of a beam stronger than the other three
    with pytest.raises(sh.StaticsError, match="material"):
        sh.lattice_cells(with_(bond_mat=m))


def test_as_generated_strengths_per_beam():
    """bodies::as_generated (restated in the oracle) makes each beam's bonds strong enough, by
    MARGIN = 2, for what the same beam of the generated lattice uses standing: afterwards no
    bond uses more than 1/2 of its strength standing (equality for a beam's most used bond where
    that need is at least 1/2), and st_lattice_cells reports the strengths beam by beam."""
    L, sites, held, mats = built(n=10, seed=4)
    mats[:, 3:] = 1.0
    L = orc.Lattice(0, sites.astype(np.int64), mats, held)
    u0 = L.used_standing(tol=1e-10)
    G = orc.Lattice(0, sites.astype(np.int64), mats, held)
    G.as_generated(L, tol=1e-10)
    u1 = G.used_standing(tol=1e-10)
    need = np.maximum(2.0 * u0.reshape(-1, 4).max(axis=1), 1.0)
    assert np.allclose(G.bond_mat[:, 3].reshape(-1, 4), need[:, None], rtol=0, atol=0)
    assert u1.max() <= 0.5 + 1e-9
    strong = need > 1.0
    assert np.allclose(u1.reshape(-1, 4).max(axis=1)[strong], 0.5, rtol=1e-9)
    c = sh.lattice_cells(G)
    bob = c["bond_index"][0::4] // 4
    s = c["strength"].reshape(-1, 3)[bob]
    assert (s[:, 0] == need).all() and (s[:, 1] == need).all() and (s[:, 2] == need).all()
