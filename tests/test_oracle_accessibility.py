"""Oracle pin of the hard-sphere accessibility grid (src/computegridsheet.h:
101-165): for one atom in a cubic cell the nearest of the 27 images is the
minimum image, so a grid point is accessible iff its minimum-image distance
is at least sigma -- a closed form independent of the image loop."""
import numpy as np
import pytest

from oracle import oracle


@pytest.mark.parametrize("L,sig,pos", [(10.0, 3.0, (0.0, 0.0, 0.0)), (7.3, 2.9, (0.31, 0.77, 0.05)),
                                       (12.0, 5.5, (0.5, 0.5, 0.5))])
def test_single_atom_cubic_closed_form(L, sig, pos):
    t = np.diag([L, L, L])
    atoms = np.array([[pos[0], pos[1], pos[2], 100.0, sig * sig]])
    n = [13, 11, 9]
    got = oracle.accessibility_grid(t, atoms, n)
    f = np.stack(np.meshgrid(*[np.arange(k) / (k - 1) for k in n], indexing="ij"), -1)
    d = f - np.array(pos)
    d -= np.round(d)
    r2 = np.sum((d * L) ** 2, axis=-1)
    assert np.array_equal(got, np.where(r2 < sig * sig, 0.0, 1.0))


def sheared_case():
    """a = (6,0,0), b = (25,1,0): at f = (0.5, 0.5, 0), i.e. (15.5, 0.5, 0),
    the nearest images with offsets in [-1, 1]^3 are 3.54 A away, while the
    image at offset (3, 0, 0), (18, 0, 0), is 1.58 A away."""
    t = np.column_stack([[6.0, 0.0, 0.0], [25.0, 1.0, 0.0], [0.0, 0.0, 20.0]])
    atoms = np.array([[0.0, 0.0, 0.0, 50.0, 9.0]])  # sigma = 3 A
    return t, atoms


def test_only_images_up_to_one_cell():
    t, atoms = sheared_case()
    got = oracle.accessibility_grid(t, atoms, [3, 3, 3])
    x = t @ np.array([0.5, 0.5, 0.0])
    near = min(np.sum((x - t @ np.array([i, j, k])) ** 2)
               for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1))
    far = np.sum((x - t @ np.array([3.0, 0.0, 0.0])) ** 2)
    assert near >= 9.0 and far < 9.0
    assert got[1, 1, 0] == 1.0  # the reference ignores the image three cells away
