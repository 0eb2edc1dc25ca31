"""The fused 3D SmoothCorner path's decomposition, pinned on the CPU oracle.

The reference smooths a 3D lattice in a fixed order (d3q15.h:199-220): the 12 edge lines in full
(corners included), then the 8 corners (d3q15.h:1241-1303).  With at least 3 nodes per axis the
fused path computes the same result from the post-boundary states of face nodes only:
* an edge node (not a corner) = 0.5 * (sum of its two inward face neighbours);
* a corner = (E_x + E_y + E_z) / 3, where E_a is the first of those means taken at the corner's
  inward neighbour along axis a (an edge node);
* every other node unchanged.
The intermediate corner values the edge lines write are overwritten by the corner step, and no
edge node reads a corner or another edge node.  These tests run the oracle's reference-order
smoothing (itself pinned to the compiled reference's fixtures) on random states and require the
closed form to match it bit for bit, so that a wrong neighbour, axis order or summation order
fails."""
import numpy as np
import pytest

from _oracle import Lattice, NC


def _closed_form(f, lx, ly, lz):
    nc = f.shape[0]
    out = f.copy()
    idx = lambda i, j, k: i + lx * (j + ly * k)
    dims = (lx, ly, lz)

    def inward(c):
        # inward neighbours along each axis on which c lies on a face, in axis order
        nb = []
        for ax in range(3):
            d = 1 if c[ax] == 0 else (-1 if c[ax] == dims[ax] - 1 else 0)
            if d:
                c2 = list(c)
                c2[ax] += d
                nb.append(tuple(c2))
        return nb

    def edge_mean(c):
        a, b = inward(c)
        return 0.5 * (f[:, idx(*a)] + f[:, idx(*b)])

    for k in range(lz):
        for j in range(ly):
            for i in range(lx):
                nb = inward((i, j, k))
                if len(nb) == 2:
                    out[:, idx(i, j, k)] = edge_mean((i, j, k))
                elif len(nb) == 3:
                    ex, ey, ez = (edge_mean(c) for c in nb)
                    out[:, idx(i, j, k)] = (ex + ey + ez) / 3.0
    return out


@pytest.mark.parametrize("dims", [(3, 3, 3), (5, 4, 6), (7, 5, 5), (4, 9, 3)])
@pytest.mark.parametrize("kind_name", ["D3Q15", "D3Q7", "D3Q27"])
def test_reference_order_smoothing_equals_closed_form(dims, kind_name):
    import _oracle

    kind = getattr(_oracle, kind_name)
    lx, ly, lz = dims
    rng = np.random.default_rng(sum(dims) + NC[kind])
    f = rng.uniform(0.0, 1.0, (NC[kind], lx * ly * lz)) * 10.0 ** rng.integers(-3, 3, (NC[kind], lx * ly * lz))
    L = Lattice(kind, lx, ly, lz)
    L.set(f)
    _oracle.lib().ol_smooth_corner(L.h)
    np.testing.assert_array_equal(L.get(), _closed_form(f, lx, ly, lz))


def test_closed_form_catches_a_wrong_corner_order():
    """A corner summed in another axis order differs in the last bits on random states, so the test
    above tells the two apart."""
    import _oracle

    lx, ly, lz = 5, 4, 6
    rng = np.random.default_rng(3)
    f = rng.uniform(0.0, 1.0, (NC[_oracle.D3Q15], lx * ly * lz)) * 10.0 ** rng.integers(-3, 3, (15, lx * ly * lz))
    L = Lattice(_oracle.D3Q15, lx, ly, lz)
    L.set(f)
    _oracle.lib().ol_smooth_corner(L.h)
    want = L.get()
    got = _closed_form(f, lx, ly, lz)
    # corner (0, 0, 0); its edge neighbours (1,0,0), (0,1,0), (0,0,1) and their face neighbours
    e = lambda a, b: 0.5 * (f[:, a] + f[:, b])
    ex = e(1 + lx, 1 + lx * ly)
    ey = e(lx + 1, lx + lx * ly)
    ez = e(lx * ly + 1, lx * ly + lx)
    np.testing.assert_array_equal(got[:, 0], want[:, 0])
    np.testing.assert_array_equal((ex + ey + ez) / 3.0, want[:, 0])
    assert np.any((ez + ey + ex) / 3.0 != want[:, 0])
