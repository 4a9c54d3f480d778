"""CPU checks that pin the oracle (no GPU).

The reference (Julia) cannot run here and its tests hold no golden vectors,
so the oracle is pinned by what the mathematics fixes: the DFT definition,
numpy's FFT, a brute-force circular sum that shares no code with the branch
recursion, and invariants of the operator (linearity, reciprocity,
translation invariance -- the identity of test/extSlfTest.jl:45-47)."""
import numpy as np
import pytest

H = 1.0 / 32


def _rx(orc, shape, seed):
    return orc.complex_gaussian(shape, seed)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 29, 30, 49, 128, 145])
def test_fft_matches_dft_definition(orc, n):
    rng = np.random.default_rng(n)
    v = rng.standard_normal((n, 3, 2)) + 1j * rng.standard_normal((n, 3, 2))
    out = orc.fft_axis(v, 0, -1)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    ref = np.einsum("mn,nab->mab", W, v)
    assert np.abs(out - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    inv = orc.fft_axis(out, 0, +1) / n
    assert np.abs(inv - v).max() <= 1e-12 * max(1.0, np.abs(v).max())


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fft_axis_matches_numpy(orc, axis):
    rng = np.random.default_rng(axis)
    v = np.asfortranarray(rng.standard_normal((6, 10, 15, 3)) + 1j * rng.standard_normal((6, 10, 15, 3)))
    out = orc.fft_axis(v, axis, -1)
    assert np.allclose(out, np.fft.fft(v, axis=axis), rtol=0, atol=1e-12)


def test_gauss_legendre_rule(orc):
    # exact for polynomials of degree 2n-1 on [0,1]
    for n in (1, 2, 4, 8, 32):
        x, w = orc.gl_rule(n)
        for p in range(2 * n):
            assert abs(np.dot(w, x ** p) - 1.0 / (p + 1)) < 1e-13


@pytest.mark.parametrize("N", [(4, 4, 4), (6, 4, 8), (5, 3, 7), (1, 2, 3), (2, 2, 2), (1, 1, 1)])
@pytest.mark.parametrize("sym", [(1, 1, 1), (0, 0, 0)])
def test_self_matvec_vs_brute_force(orc, N, sym):
    toe = orc.toe_model(N, (H, H, H), model=0)
    crc = orc.crc_self(N, toe)
    fur = orc.fur_from_crc(N, N, sym, crc)
    x = _rx(orc, N + (3,), 7)
    y = orc.matvec(N, N, sym, fur, x)
    yb = orc.brute(N, N, crc, x)
    assert orc.rel_l2(y, yb) < 1e-13


@pytest.mark.parametrize("S,T", [((4, 6, 2), (6, 8, 4)), ((3, 5, 2), (5, 3, 6)), ((8, 2, 2), (2, 8, 4)),
                                 ((4, 4, 4), (4, 4, 4))])
def test_external_matvec_vs_brute_force(orc, S, T):
    crc = orc.crc_ext(S, T, (H, H, H), (0.3, -0.2, 0.5))
    fur = orc.fur_from_crc(S, T, (0, 0, 0), crc)
    x = _rx(orc, tuple(S) + (3,), 11)
    assert orc.rel_l2(orc.matvec(S, T, (0, 0, 0), fur, x), orc.brute(S, T, crc, x)) < 1e-13


def test_brute_force_is_direct_block_toeplitz(orc):
    # the brute-force circular sum equals the block-Toeplitz sum written
    # directly from the real-space tensor with reflection signs
    N = (3, 2, 4)
    toe = orc.toe_model(N, (H, H, H))
    crc = orc.crc_self(N, toe)
    x = _rx(orc, N + (3,), 3)
    y = np.zeros(N + (3,), complex)
    for t in np.ndindex(*N):
        for s in np.ndindex(*N):
            d = [a - b for a, b in zip(t, s)]
            G = toe[:, :, abs(d[0]), abs(d[1]), abs(d[2])].copy()
            f = [1 if v >= 0 else -1 for v in d]
            for i in range(3):
                for j in range(3):
                    a, b = min(i, j), max(i, j)
                    y[t + (i,)] += G[a, b] * (1 if i == j else f[i] * f[j]) * x[s + (j,)]
    assert orc.rel_l2(orc.brute(N, N, crc, x), y) < 1e-14


def test_fur_mirror_symmetry(orc):
    # G^(L-m) = sigma G^(m) per axis: symmetric (truncated) and full storage
    # give the same operator
    N = (6, 6, 6)
    fur_s, crc = orc.self_operator_fur(N)
    fur_f = orc.fur_from_crc(N, N, (0, 0, 0), crc)
    x = _rx(orc, N + (3,), 5)
    assert orc.rel_l2(orc.matvec(N, N, (1, 1, 1), fur_s, x), orc.matvec(N, N, (0, 0, 0), fur_f, x)) < 1e-14


def test_linearity(orc):
    N = (4, 6, 4)
    fur, _ = orc.self_operator_fur(N)
    x1, x2 = _rx(orc, N + (3,), 1), _rx(orc, N + (3,), 2)
    a, b = 0.3 - 1.2j, -2.0 + 0.5j
    lhs = orc.matvec(N, N, (1, 1, 1), fur, a * x1 + b * x2)
    rhs = a * orc.matvec(N, N, (1, 1, 1), fur, x1) + b * orc.matvec(N, N, (1, 1, 1), fur, x2)
    assert orc.rel_l2(lhs, rhs) < 1e-13


def test_reciprocity(orc):
    # the self operator is complex symmetric: x^T (G y) = y^T (G x)
    N = (5, 4, 6)
    fur, _ = orc.self_operator_fur(N)
    x, y = _rx(orc, N + (3,), 3), _rx(orc, N + (3,), 4)
    a = np.sum(x * orc.matvec(N, N, (1, 1, 1), fur, y))
    b = np.sum(y * orc.matvec(N, N, (1, 1, 1), fur, x))
    assert abs(a - b) / abs(a) < 1e-13


@pytest.mark.parametrize("model", [0, 1])
def test_translation_invariance_self_vs_external(orc, model):
    # test/extSlfTest.jl:45-47: a self operator on the union U of two touching
    # volumes A, B restricted to B equals the external operator A -> B
    A = (4, 4, 4)
    U = (4, 8, 4)
    fu, _ = orc.self_operator_fur(U, model=model, qf=3, qn=5)
    xu = np.zeros(U + (3,), complex, order="F")
    xa = _rx(orc, A + (3,), 9)
    xu[:, :4, :, :] = xa
    yu = orc.matvec(U, U, (1, 1, 1), fu, xu)
    crc = orc.crc_ext(A, A, (H, H, H), (0.0, 4 * H, 0.0), model=model, qf=3, qn=5)
    fe = orc.fur_from_crc(A, A, (1, 0, 1), crc)
    ye = orc.matvec(A, A, (1, 0, 1), fe, xa)
    assert orc.rel_l2(ye, yu[:, 4:, :, :]) < 1e-12


def test_fp32_restatement_tracks_fp64(orc):
    N = (8, 8, 8)
    fur, _ = orc.self_operator_fur(N)
    x = _rx(orc, N + (3,), 1)
    y64 = orc.matvec(N, N, (1, 1, 1), fur, x)
    y32 = orc.matvec(N, N, (1, 1, 1), fur.astype(np.complex64), x.astype(np.complex64), dtype=np.complex64)
    assert orc.rel_l2(y32, y64) < 1e-5


def test_nonfinite_green_data_is_rejected(orc):
    N = (2, 2, 2)
    toe = orc.toe_model(N, (H, H, H))
    crc = orc.crc_self(N, toe)
    crc[0, 0, 1, 0, 0] = np.nan
    with pytest.raises(FloatingPointError):
        orc.fur_from_crc(N, N, (1, 1, 1), crc)


def test_seeded_gaussian_is_portable(orc):
    # SplitMix64 reference values (seed 0): first output 0xE220A8397B1DCDAF
    assert int(orc.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    g = orc.complex_gaussian((64, 64, 3), seed=3)
    assert abs(np.mean(np.abs(g) ** 2) - 1.0) < 0.02
    assert np.array_equal(g, orc.complex_gaussian((64, 64, 3), seed=3))


def _analytic_dyadic(sep):
    """(k^2 + grad grad) e^{ikr}/(4 pi r), k = 2 pi (test/anaTest.jl:36-44)."""
    d = np.linalg.norm(sep)
    s = 2 * np.pi * d
    sh = np.asarray(sep) / d
    k2g = 2 * np.pi ** 2 * np.exp(1j * s) / s
    return k2g * ((1 + (1j * s - 1) / s ** 2) * np.eye(3) - (1 + 3 * (1j * s - 1) / s ** 2) * np.outer(sh, sh))


@pytest.mark.parametrize("h", [1.0 / 32, 1.0 / 64])
def test_facepair_far_field_matches_analytic_dyadic(orc, h):
    # G / V tends to the point dyadic, up to the cube average of the phase,
    # a factor 1 - (k h)^2 / 12 for source and target cells together
    up = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    kh2 = (2 * np.pi * h) ** 2 / 12
    cells = [(8, 5, 3), (16, 10, 6), (32, 20, 12)]
    scale = int(round(1 / (32 * h)))
    N = tuple(max(c[a] for c in cells) * scale + 1 for a in range(3))
    toe = orc.toe_model(N, (h, h, h), model=2, qn=10)
    for c in cells:
        r = tuple(v * scale for v in c)          # same physical separation for both h
        G = toe[:, :, r[0], r[1], r[2]] / h ** 3
        ana = _analytic_dyadic(np.array(r) * h)
        err = max(abs(G[i, j] - ana[i, j]) for i, j in up) / np.abs(ana).max()
        assert 0.5 * kh2 < err < 1.5 * kh2, (h, r, err, kh2)


def test_facepair_components_and_signs(orc):
    # diagonal/off-diagonal structure of the face-pair tensor along an axis
    # and a diagonal: on the x axis only xx, yy=zz are non-zero; on the
    # (1,1,0) diagonal xy is non-zero and xz = yz = 0
    h = 1.0 / 32
    toe = orc.toe_model((12, 12, 2), (h, h, h), model=2, qn=10)
    g = toe[:, :, 10, 0, 0]
    assert abs(g[0, 1]) < 1e-12 * abs(g[0, 0]) and abs(g[0, 2]) < 1e-12 * abs(g[0, 0])
    assert abs(g[1, 1] - g[2, 2]) < 1e-10 * abs(g[1, 1])
    g = toe[:, :, 7, 7, 0]
    assert abs(g[0, 1]) > 1e-3 * abs(g[0, 0])
    assert abs(g[0, 2]) < 1e-12 * abs(g[0, 1]) and abs(g[1, 2]) < 1e-12 * abs(g[0, 1])
    assert abs(g[0, 0] - g[1, 1]) < 1e-10 * abs(g[0, 0])


def test_golden_fixtures(orc):
    # regression pins written by tests/golden/make_golden.py
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = np.load(os.path.join(here, "self_8.npz"))
    N = tuple(int(v) for v in g["cells"])
    toe = orc.toe_model(N, (float(g["h"]),) * 3, model=2, qn=10)
    assert np.abs(toe - g["toe"]).max() <= 1e-12 * np.abs(g["toe"]).max()
    fur = orc.fur_from_crc(N, N, (1, 1, 1), orc.crc_self(N, toe))
    assert np.abs(fur - g["fur"]).max() <= 1e-12 * np.abs(g["fur"]).max()
    assert orc.rel_l2(orc.matvec(N, N, (1, 1, 1), g["fur"], g["x"]), g["y"]) < 1e-14
    e = np.load(os.path.join(here, "ext_4x6x2_6x8x4.npz"))
    S, T = tuple(int(v) for v in e["src"]), tuple(int(v) for v in e["trg"])
    assert orc.rel_l2(orc.matvec(S, T, (0, 0, 0), e["fur"], e["x"]), e["y"]) < 1e-14
    crc = orc.crc_ext(S, T, (float(e["h"]),) * 3, tuple(e["delta"]), model=0)
    assert orc.rel_l2(orc.brute(S, T, crc, e["x"]), e["y"]) < 1e-13
