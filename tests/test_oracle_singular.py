"""CPU pins of the oracle's restatement of the reference's singular face-pair
integrals (DIRECTFN wekS / wekE / wekV, the closed-form static panel integrals,
the egoFunSng! correction tables) and of its regular face pairs, against the
independent difference-coordinate quadrature of tests/panel_quadrature.py and
against what the mathematics fixes.

Tolerances: the static closed forms are exact (1e-12); the singular values at
the reference's default order 32 carry the reductions' own quadrature error
(measured about 1e-11 for self and edge pairs, 7e-10 for vertex pairs, whose
full-kernel reduction is not centred on the shared vertex); the assembled
tensor at touching offsets inherits that error amplified by the cancellation
between faces (a few 1e-9); regular far pairs are converged to 1e-12."""
import numpy as np
import pytest

import panel_quadrature as pq

H32 = (1 / 32,) * 3
ANISO = (1 / 32, 1 / 24, 1 / 40)


def test_coincident_square_closed_form_is_the_textbook_value(orc):
    # the integral of 1/|r - r'| over the unit square with itself is
    # 4 ln(1 + sqrt 2) - (4/3)(sqrt 2 - 1); rSrfSlf carries 1/(4 pi)
    exact = 4 * np.log(1 + np.sqrt(2)) - 4 / 3 * (np.sqrt(2) - 1)
    assert abs(orc.r_srf(0, 1.0, 1.0).real * 4 * np.pi - exact) < 1e-13 * exact


@pytest.mark.parametrize("la,lb,lc", [(1.0, 1.0, 1.0), (1.0, 2.0, 3.0), (0.5, 1.3, 0.7), (1 / 32, 1 / 24, 1 / 40)])
def test_static_closed_forms_match_independent_quadrature(orc, la, lb, lc):
    slf = pq.parallel_pair(((0, la), (0, lb)), ((0, la), (0, lb)), 0.0, 1.0, kernel="static")
    flt = pq.parallel_pair(((0, la), (0, lb)), ((0, la), (lb, 2 * lb)), 0.0, 1.0, kernel="static")
    # right angle along the shared edge la: one rectangle spans lb, the other lc
    crn = pq.perp_pair((0, la), (0, lb), 0.0, (0, la), 0.0, (0, lc), 1.0, kernel="static")
    for which, ref in ((0, slf), (1, flt), (2, crn)):
        val = orc.r_srf(which, la, lb, lc)
        assert abs(val - ref) <= 1e-12 * abs(ref), (which, val, ref)


def _exact_singular(s, zeta, n=32):
    """The 18 panel pairs that wS / wE / wV stand for, by geometry."""
    sx, sy, sz = s
    pp, pe = pq.parallel_pair, pq.perp_pair
    S = [pp(((0, sy), (0, sz)), ((0, sy), (0, sz)), 0, zeta, n),
         pp(((0, sx), (0, sz)), ((0, sx), (0, sz)), 0, zeta, n),
         pp(((0, sx), (0, sy)), ((0, sx), (0, sy)), 0, zeta, n)]
    # flat edges: [xxY xxZ yyX yyZ zzX zzY]; cornered: [xy xz yz] share the
    # z, y and x edges
    E = [pp(((0, sy), (0, sz)), ((sy, 2 * sy), (0, sz)), 0, zeta, n),
         pp(((0, sy), (0, sz)), ((0, sy), (sz, 2 * sz)), 0, zeta, n),
         pp(((0, sx), (0, sz)), ((sx, 2 * sx), (0, sz)), 0, zeta, n),
         pp(((0, sx), (0, sz)), ((0, sx), (sz, 2 * sz)), 0, zeta, n),
         pp(((0, sx), (0, sy)), ((sx, 2 * sx), (0, sy)), 0, zeta, n),
         pp(((0, sx), (0, sy)), ((0, sx), (sy, 2 * sy)), 0, zeta, n),
         pe((0, sz), (0, sy), 0.0, (0, sz), 0.0, (0, sx), zeta, n),
         pe((0, sy), (0, sz), 0.0, (0, sy), 0.0, (0, sx), zeta, n),
         pe((0, sx), (0, sz), 0.0, (0, sx), 0.0, (0, sy), zeta, n)]
    # vertices: coplanar diagonal neighbours [xx yy zz], then perpendicular
    # faces meeting at one corner [xy xz yz]
    V = [pp(((0, sy), (0, sz)), ((sy, 2 * sy), (sz, 2 * sz)), 0, zeta, n),
         pp(((0, sx), (0, sz)), ((sx, 2 * sx), (sz, 2 * sz)), 0, zeta, n),
         pp(((0, sx), (0, sy)), ((sx, 2 * sx), (sy, 2 * sy)), 0, zeta, n),
         pe((0, sz), (0, sy), 0.0, (sz, 2 * sz), sy, (0, sx), zeta, n),
         pe((0, sy), (0, sz), 0.0, (sy, 2 * sy), sz, (0, sx), zeta, n),
         pe((0, sx), (0, sz), 0.0, (sx, 2 * sx), sz, (0, sy), zeta, n)]
    return np.array(S), np.array(E), np.array(V)


@pytest.mark.parametrize("s", [H32, ANISO])
def test_singular_values_match_independent_quadrature(orc, s):
    S, E, V = _exact_singular(s, 1.0)
    for kind, ex, tol in (("S", S, 1e-10), ("E", E, 1e-10), ("V", V, 2e-9)):
        w = orc.wek(kind, s, 1.0, 32)
        err = np.abs(w - ex) / np.abs(ex)
        assert err.max() <= tol, (kind, err)


def test_singular_values_converge_with_order(orc):
    # test/intConTest.jl: order intOrd against intOrd + 8 below 1e-8
    for kind in "SEV":
        a = orc.wek(kind, H32, 1.0, 32)
        b = orc.wek(kind, H32, 1.0, 40)
        assert (np.abs(a - b) / np.abs(b)).max() < 1e-8, kind


def test_independent_quadrature_is_converged(orc):
    S, E, V = _exact_singular(ANISO, 1.0, n=24)
    S2, E2, V2 = _exact_singular(ANISO, 1.0, n=32)
    for a, b in ((S, S2), (E, E2), (V, V2)):
        assert (np.abs(a - b) / np.abs(b)).max() < 1e-13


@pytest.mark.parametrize("s", [H32, ANISO])
def test_touching_offsets_match_independent_face_pairs(orc, s):
    # the whole tensor at the 8 offsets egoFunSng! corrects, from all 36 face
    # pairs integrated independently: pins the mask / corVal tables (a value
    # put on the wrong face pair shows at O(1) on the anisotropic cell) and
    # the regular pairs next to them
    toe = orc.toe_model((2, 2, 2), s, model=3)
    for idx in np.ndindex(2, 2, 2):
        sep = [idx[d] * s[d] for d in range(3)]
        G = pq.green_tensor(sep, s, 1.0)
        err = np.abs(G - toe[:, :, idx[0], idx[1], idx[2]]).max() / np.abs(G).max()
        tol = 1e-10 if sum(idx) <= 1 else 1e-8
        assert err <= tol, (idx, err)


def test_self_cell_value(orc):
    # G_xx(0) at h = 1/32, frqPhz = 1: -0.32950946120 (singular values plus
    # the identity term); a plain product rule on touching faces gives -0.3797
    toe = orc.toe_model((1, 1, 1), H32, model=3)
    G = pq.green_tensor((0.0, 0.0, 0.0), H32, 1.0)
    assert abs(toe[0, 0, 0, 0, 0] - G[0, 0]) < 1e-11
    assert abs(toe[0, 0, 0, 0, 0].real + 0.3295094612) < 1e-9
    for i in range(3):
        for j in range(3):
            if i != j:
                assert abs(toe[i, j, 0, 0, 0]) < 1e-14


@pytest.mark.parametrize("idx", [(2, 0, 0), (3, 1, 2), (5, 4, 0)])
def test_regular_offsets_match_independent_face_pairs(orc, idx):
    toe = orc.toe_model((6, 5, 3), H32, model=3)
    sep = [idx[d] / 32 for d in range(3)]
    G = pq.green_tensor(sep, H32, 1.0, n=24)
    assert np.abs(G - toe[:, :, idx[0], idx[1], idx[2]]).max() <= 1e-11 * np.abs(G).max()


def test_regular_pairs_at_a_coarse_cell(orc):
    # h = 1/8: the phase varies by ~0.8 rad across a cell, so the regular
    # rules need the phase-dependent order, not only the distance one
    s = (1 / 8,) * 3
    toe = orc.toe_model((4, 1, 1), s, model=3)
    for k in (2, 3):
        G = pq.green_tensor((k / 8, 0.0, 0.0), s, 1.0, n=24)
        assert np.abs(G - toe[:, :, k, 0, 0]).max() <= 1e-11 * np.abs(G).max()


def test_external_contact_uses_reflected_self_values(orc):
    # a target cell one cell below the source along y (contact): the value is
    # the self tensor at the absolute offset with the y-flips (egoCntOut!
    # through the contact volume's circulant, src/glaGen.jl:250-304)
    s = H32
    crc = orc.crc_ext((1, 1, 1), (1, 1, 1), s, (0.0, -s[1], 0.0), model=3)
    G = pq.green_tensor((0.0, -s[1], 0.0), s, 1.0)
    assert np.abs(crc[:, :, 0, 0, 0] - G).max() <= 1e-10 * np.abs(G).max()
