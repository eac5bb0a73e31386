"""st_lattice_cells (host only, no GPU): a lattice as the reference holds it back to the cells
lattice::build made it from, against the cells the oracle built it from; refusal of lattices
the device's operator cannot take (gathered ones, broken beam geometry); per-beam strengths
of bodies::as_generated lattices."""
import numpy as np
import pytest

import oracle as orc
import statics_hip as sh
from harness import FACES, PAIR_OF, mats_from


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
    bob = c["beam_of_bond"]
    cell_dev = bob // 15
    assert (bob[0::4] == bob[1::4]).all() and (bob[0::4] == bob[3::4]).all()
    for b in range(0, L.n_bonds, 4 * 7):
        i_or = L.beam_cell[b // 4]
        assert cell_of_site[tuple(c["sites"][cell_dev[b]].tolist())] == i_or
    # held-held beams are absent, the others present
    nbeams = np.unpackbits(c["beams"].view(np.uint8)).sum()
    assert nbeams == L.n_bonds // 4


def test_gathered_lattice_is_refused():
    L, *_ = built(n=12)
    G = L.gathered()
    with pytest.raises(sh.StaticsError) as e:
        sh.lattice_cells(G)
    assert e.value.code == sh.ST_BAD_INPUT and "section" in str(e.value)


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
    bob = c["beam_of_bond"][0::4]
    s = c["strength"].reshape(-1, 3)[bob]
    assert (s[:, 0] == need).all() and (s[:, 1] == need).all() and (s[:, 2] == need).all()
