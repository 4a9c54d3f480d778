"""Physics check of the reference's Green-tensor construction as the oracle
restates it (model 3): test/posDefTest.jl -- the anti-Hermitian part of the
dense self operator, (i/2)(conj(zeta)^2 G^H - zeta^2 G), has no eigenvalue
below -1e-6 (the operator of a lossless medium is passive)."""
import numpy as np


def test_anti_hermitian_part_is_positive_semidefinite(orc):
    N = (4, 4, 4)
    h = 1 / 32
    fur, _ = orc.self_operator_fur(N, h=h, model=3)
    n = 3 * int(np.prod(N))
    X = np.eye(n, dtype=np.complex128).reshape(N + (3, n), order="F")
    G = orc.matvec(N, N, (1, 1, 1), fur, X).reshape(n, n, order="F")
    zeta = 1.0
    asym = 0.5j * np.conj(zeta) ** 2 * G.conj().T - 0.5j * zeta ** 2 * G
    ev = np.linalg.eigvalsh(0.5 * (asym + asym.conj().T))
    assert ev.min() > -1e-6, ev.min()
    # and the construction is reciprocal: G is complex symmetric
    assert np.abs(G - G.T).max() <= 1e-12 * np.abs(G).max()
