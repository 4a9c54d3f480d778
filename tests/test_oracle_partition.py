"""The oracle's partitioned (multi-resolution) external operator.

tensor_mixed (cells of different sizes) is pinned by subdivision: the face-pair
sum of srfSum! over a cell is additive over a split of that cell (the faces
inside cancel in pairs, the outer faces split into pieces), so the tensor
between a coarse cell and a fine one equals a sum of equal-size model-3
tensors over the coarse cell's fine subcells, each face pair weighted by the
ratio of the srfScl factors (src/glaInt.jl:284-305) of the two
parametrisations.  For cells of one shape that ratio is 1 / 8 (coarse target)
or 1 (coarse source) for every component; for a target twice as long on x
only it is 1 / 2 for xx, yy, zz, yz and 1 for xy, xz -- the trace of srfScl
indexing its table with the target face where the source face belongs.

The partition matvec (genPrt / leaf sums / mrgPrt) is pinned against a direct
double sum over all (target cell, source cell) pairs of the whole volumes,
with cell centres from the volume definitions (GlaVol, src/glaMem.jl:136-147)
and the tensor's upper triangle as the matvec applies it."""
import itertools

import numpy as np
import pytest

h = 1.0 / 32


def _sub_sum(orc, sep, coarse_t, scale):
    """Sum of model-3 tensors between the fine subcells of the coarse cell
    (target when coarse_t, else source) and the fine other cell."""
    acc = np.zeros((3, 3), complex)
    for o in itertools.product((-0.5, 0.5), repeat=3):
        d = [o[a] * h * (scale[a] == 2) for a in range(3)]
        if all(scale[a] == 1 for a in range(3)):
            raise ValueError
        if any(scale[a] == 1 and o[a] > 0 for a in range(3)):
            continue        # axes that are not split carry one subcell
        s = [sep[a] + (d[a] if coarse_t else -d[a]) for a in range(3)]
        acc += orc.tensor_ref(s, (h, h, h))
    return acc


def test_mixed_equal_cells_is_model3(orc):
    for sep in [(3 * h, 2 * h, h), (-5 * h, 0.0, 2 * h), (7 * h, -3 * h, -4 * h)]:
        a = orc.tensor_mixed(sep, (h, h, h), (h, h, h))
        b = orc.tensor_ref(sep, (h, h, h))
        assert np.array_equal(a, b)


@pytest.mark.parametrize("sep", [(3.5 * h, 0.5 * h, 1.5 * h), (6.5 * h, 3.5 * h, -2.5 * h),
                                 (-4.5 * h, 1.5 * h, 0.5 * h)])
def test_mixed_subdivision_same_shape(orc, sep):
    a = orc.tensor_mixed(sep, (2 * h,) * 3, (h,) * 3)
    ref = _sub_sum(orc, sep, True, (2, 2, 2)) / 8
    assert np.abs(a - ref).max() <= 1e-11 * np.abs(ref).max()
    a = orc.tensor_mixed(sep, (h,) * 3, (2 * h,) * 3)
    ref = _sub_sum(orc, sep, False, (2, 2, 2))
    assert np.abs(a - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("sep", [(3.5 * h, h, 2 * h), (-5.5 * h, -2 * h, h)])
def test_mixed_subdivision_target_long_on_x(orc, sep):
    a = orc.tensor_mixed(sep, (2 * h, h, h), (h, h, h))
    sub = _sub_sum(orc, sep, True, (2, 1, 1))
    w = np.full((3, 3), 0.5)
    w[0, 1] = w[1, 0] = w[0, 2] = w[2, 0] = 1.0
    ref = w * sub
    # the upper triangle is what the operator stores
    iu = np.triu_indices(3)
    assert np.abs(a[iu] - ref[iu]).max() <= 1e-11 * np.abs(ref).max()
    # a dropped srfScl swap would give 1/2 on xy: the check separates them
    assert np.abs(a[0, 1] - 0.5 * sub[0, 1]) > 1e-3 * np.abs(sub[0, 1])


def _centres(cel, scl, org):
    return [org[a] - scl[a] * (cel[a] - 1) / 2 + scl[a] * np.arange(cel[a]) for a in range(3)]


def _direct(orc, Tf, sT, orgT, Sf, sS, orgS, x):
    ct, cs = _centres(Tf, sT, orgT), _centres(Sf, sS, orgS)
    y = np.zeros(tuple(Tf) + (3,), complex, order="F")
    for it in itertools.product(*[range(n) for n in Tf]):
        rt = [ct[a][it[a]] for a in range(3)]
        for js in itertools.product(*[range(n) for n in Sf]):
            g = orc.tensor_mixed([rt[a] - cs[a][js[a]] for a in range(3)], sT, sS)
            gs = np.triu(g) + np.triu(g, 1).T
            y[it] += gs @ x[js]
    return y


@pytest.mark.parametrize("case", ["src_div_x", "trg_div_y", "both"])
def test_partition_matvec_equals_direct_sum(orc, case):
    if case == "src_div_x":
        sT, sS, Tf, Sf = (2 * h, h, h), (h, h, h), (2, 2, 2), (4, 2, 2)
    elif case == "trg_div_y":
        sT, sS, Tf, Sf = (h, h, h), (h, 2 * h, h), (2, 4, 2), (2, 2, 2)
    else:
        sT, sS, Tf, Sf = (2 * h, h, h), (h, 2 * h, h), (2, 4, 2), (4, 2, 2)
    orgT, orgS = (0.0, 0.0, 0.0), (9 * h, -2 * h, 3 * h)
    fur, S, T, sdiv, tdiv = orc.partition_fur(Tf, Sf, sT, sS, orgT, orgS)
    assert tuple(s * d for s, d in zip(S, sdiv)) == Sf and tuple(t * d for t, d in zip(T, tdiv)) == Tf
    x = orc.complex_gaussian(tuple(Sf) + (3,), seed=77)
    y = orc.partition_matvec(S, T, sdiv, tdiv, fur, x)
    yd = _direct(orc, Tf, sT, orgT, Sf, sS, orgS, x)
    assert orc.rel_l2(y, yd) <= 1e-12
