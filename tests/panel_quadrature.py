"""Independent quadrature of face-pair integrals of the scalar Green function,
used by the tests to pin the oracle's singular and regular face-pair values.

For two axis-aligned rectangles the 4-D integral of g(|p_t - p_s|) reduces,
in difference coordinates, to a lower-dimensional integral whose weight is the
overlap length of the intervals along each shared in-plane axis (a piecewise
linear function): 2-D for parallel rectangles, 3-D for perpendicular ones.
The pieces between the kinks of those weights are integrated with tensor
Gauss-Legendre rules; a piece with the singular point (zero separation) at a
corner is integrated in polar (2-D) or Duffy-pyramid (3-D) coordinates about
that corner, where r*g(r) is smooth.  Nothing here shares code with the
oracle's DIRECTFN restatement or its tensor rules.
"""
import numpy as np
from numpy.polynomial.legendre import leggauss

EPS = 1e-15


def _gl(n, a, b):
    x, w = leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (a + b), 0.5 * (b - a) * w


def _k_static(r, zeta):
    return 1.0 / (4 * np.pi * r * zeta ** 2)


def _rk_static(r, zeta):
    return np.full(np.shape(r), 1.0 / (4 * np.pi * zeta ** 2), dtype=complex)


def _k_full(r, zeta):
    return np.exp(2j * np.pi * r * zeta) / (4 * np.pi * r * zeta ** 2)


def _rk_full(r, zeta):
    return np.exp(2j * np.pi * r * zeta) / (4 * np.pi * zeta ** 2)


KERNELS = {"full": (_k_full, _rk_full), "static": (_k_static, _rk_static)}


def _overlap_pieces(p1, q1, p2, q2):
    """c(w) = |[p1,q1] ∩ [p2-w, q2-w]| as linear pieces (w0, w1, c(w0), c(w1))."""
    ks = sorted({p2 - q1, p2 - p1, q2 - q1, q2 - p1})

    def c(w):
        return max(0.0, min(q1, q2 - w) - max(p1, p2 - w))

    return [(a, b, c(a), c(b)) for a, b in zip(ks[:-1], ks[1:]) if b - a > EPS]


def _lin(w, a, b, ca, cb):
    return ca + (cb - ca) * (w - a) / (b - a)


def _at_corner(a, b):
    return a <= EPS and b >= -EPS and (abs(a) < EPS or abs(b) < EPS)


def parallel_pair(r1, r2, d, zeta, n=32, kernel="full"):
    """Integral over rectangle r1 x rectangle r2 (each ((u0,u1),(v0,v1)) in
    parallel planes a distance d apart) of g(|p2 - p1|)."""
    k, rk = KERNELS[kernel]
    tot = 0.0
    for (a, b, ca, cb) in _overlap_pieces(r1[0][0], r1[0][1], r2[0][0], r2[0][1]):
        for (c, e, cc, ce) in _overlap_pieces(r1[1][0], r1[1][1], r2[1][0], r2[1][1]):
            if abs(d) < EPS and _at_corner(a, b) and _at_corner(c, e):
                su = 1.0 if abs(a) < EPS else -1.0
                sv = 1.0 if abs(c) < EPS else -1.0
                A, B = b - a, e - c
                th0 = np.arctan2(B, A)
                for (ta, tb, rmax) in ((0.0, th0, lambda t: A / np.cos(t)), (th0, np.pi / 2, lambda t: B / np.sin(t))):
                    t, wt = _gl(n, ta, tb)
                    for ti, wti in zip(t, wt):
                        r, wr = _gl(n, 0.0, rmax(ti))
                        w1, w2 = su * r * np.cos(ti), sv * r * np.sin(ti)
                        tot += wti * np.sum(wr * rk(r, zeta) * _lin(w1, a, b, ca, cb) * _lin(w2, c, e, cc, ce))
            else:
                assert not (abs(d) < EPS and a < -EPS and b > EPS and c < -EPS and e > EPS)
                u, wu = _gl(n, a, b)
                v, wv = _gl(n, c, e)
                U, V = np.meshgrid(u, v, indexing="ij")
                R = np.sqrt(U ** 2 + V ** 2 + d * d)
                tot += np.sum(np.outer(wu, wv) * k(R, zeta) * _lin(U, a, b, ca, cb) * _lin(V, c, e, cc, ce))
    return tot


def _box_corner(L, f, n, rk, zeta):
    """Integral over the box [0,L0]x[0,L1]x[0,L2] of f(p) g(|p|), the
    singular point at the origin corner: three Duffy pyramids."""
    tot = 0.0
    t, wt = _gl(n, 0.0, 1.0)
    for ax in range(3):
        o = [i for i in range(3) if i != ax]
        s1, w1 = _gl(n, 0.0, L[o[0]])
        s2, w2 = _gl(n, 0.0, L[o[1]])
        T, S1, S2 = np.meshgrid(t, s1, s2, indexing="ij")
        W = wt[:, None, None] * w1[None, :, None] * w2[None, None, :]
        P = [None] * 3
        P[ax], P[o[0]], P[o[1]] = T * L[ax], T * S1, T * S2
        rho = np.sqrt(P[0] ** 2 + P[1] ** 2 + P[2] ** 2)
        # g(rho) rho^2 dt ... = rk(rho) * T * L[ax] / |base point|
        tot += np.sum(W * f(P) * rk(rho, zeta) * T * T * L[ax] / rho)
    return tot


def perp_pair(x1, y1, z1, x2, y2, z2, zeta, n=24, kernel="full"):
    """Rectangle 1 in the plane z = z1 spanning x in x1, y in y1; rectangle 2
    in the plane y = y2 spanning x in x2, z in z2: integral of g(|p2 - p1|)."""
    k, rk = KERNELS[kernel]
    tot = 0.0
    ua, ub = y2 - y1[1], y2 - y1[0]
    va, vb = z2[0] - z1, z2[1] - z1
    for (a, b, ca, cb) in _overlap_pieces(x1[0], x1[1], x2[0], x2[1]):
        if _at_corner(a, b) and _at_corner(ua, ub) and _at_corner(va, vb):
            sw = 1.0 if abs(a) < EPS else -1.0
            tot += _box_corner((b - a, ub - ua, vb - va), lambda P: _lin(sw * P[0], a, b, ca, cb), n, rk, zeta)
        else:
            assert not (a < -EPS < EPS < b and ua < -EPS < EPS < ub and va < -EPS < EPS < vb)
            w, ww = _gl(n, a, b)
            u, wu = _gl(n, ua, ub)
            v, wv = _gl(n, va, vb)
            Wg, Ug, Vg = np.meshgrid(w, u, v, indexing="ij")
            W = ww[:, None, None] * wu[None, :, None] * wv[None, None, :]
            tot += np.sum(W * _lin(Wg, a, b, ca, cb) * k(np.sqrt(Wg ** 2 + Ug ** 2 + Vg ** 2), zeta))
    return tot


def face(s, f):
    """Face f of a cell of size s centred at the origin (cubFac numbering:
    0/1 lower/upper x, 2/3 y, 4/5 z): (normal axis, position, ranges of the
    three axes with the normal one collapsed)."""
    n = f // 2
    pos = (0.5 if f % 2 else -0.5) * s[n]
    rng = [(-0.5 * s[d], 0.5 * s[d]) for d in range(3)]
    rng[n] = (pos, pos)
    return n, pos, rng


def face_pair(ft, fs, sep, s, zeta, n=32, kernel="full"):
    """Integral of g over target face ft (cell centred at sep) times source
    face fs (cell at the origin)."""
    nt, pt, rt = face(s, ft)
    ns, ps, rs = face(s, fs)
    rt = [(lo + sep[d], hi + sep[d]) for d, (lo, hi) in enumerate(rt)]
    if nt == ns:
        a, b = [d for d in range(3) if d != nt]
        return parallel_pair((rs[a], rs[b]), (rt[a], rt[b]), (pt + sep[nt]) - ps, zeta, n, kernel)
    c = 3 - nt - ns
    # rectangle 1: the source face (plane ns), spanning the common axis c and
    # the target's normal axis nt; rectangle 2: the target face (plane nt),
    # spanning c and ns
    return perp_pair(rs[c], rs[nt], ps, rt[c], pt + sep[nt], rt[ns], zeta, max(16, 3 * n // 4), kernel)


def green_tensor(sep, s, zeta, n=32):
    """The face-pair Green tensor at separation sep (cell-volume normalised):
    A[i][j] collects the pairs with target normal i and source normal j with
    sign (lower face +, upper face -) of each face, negated for i != j; the
    diagonal entry of axis i sums the two other axes' parallel pairs."""
    V = s[0] * s[1] * s[2]
    A = np.zeros((3, 3), complex)
    for ft in range(6):
        for fs in range(6):
            i, j = ft // 2, fs // 2
            sg = (1 if ft % 2 == 0 else -1) * (1 if fs % 2 == 0 else -1) * (1 if i == j else -1)
            A[i, j] += sg * face_pair(ft, fs, sep, s, zeta, n) / V
    G = np.zeros((3, 3), complex)
    for i in range(3):
        G[i, i] = sum(A[m, m] for m in range(3) if m != i)
        for j in range(3):
            if i != j:
                G[i, j] = A[i, j]
    if not any(sep):
        G -= np.eye(3) / zeta ** 2
    return G
