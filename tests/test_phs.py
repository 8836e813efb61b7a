"""P3: the prolate-hyperspheroid (PHS) transform of the optimise passes
(pop_planner.cu:289-346), pinned by the geometry it must have rather than by
a restatement of its formula.

For foci f1 (start), f2 (goal 0) at distance c_min and best cost c_best the
transform TF = R * diag(c_best/2, r, ..., r), r = sqrt(c_best^2 - c_min^2)/2,
with centre C = (f1 + f2)/2 must satisfy:
  * R orthogonal with R e_1 = (f2 - f1) / c_min (the transverse axis);
  * column 0 of TF of length c_best/2, every other column of length r;
  * every point C + TF z with |z| = 1 on the hyperspheroid's boundary,
    |x - f1| + |x - f2| = c_best, and every |z| < 1 inside it;
  * degenerate cases: coincident foci give a ball of radius c_best/2
    (R = I), and c_best <= c_min collapses the conjugate radii to 0.
A wrong Householder vector, swapped radii, a missing factor 1/2 or a centre
off the midpoint each fail one of these.  The GPU tests check that the
device's transform equals this host function's and that unclamped device
samples lie inside the hyperspheroid.
"""
import ctypes

import numpy as np
import pytest

from mrpop import _lib


def phs(f1, f2, cb):
    D = len(f1)
    L = _lib.lib()
    fn = L.mrpop_diag_phs_transform
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    f1 = np.ascontiguousarray(f1, np.float32)
    f2 = np.ascontiguousarray(f2, np.float32)
    tf = np.zeros((D, D), np.float32)
    c = np.zeros(D, np.float32)
    assert fn(D, f1.ctypes.data, f2.ctypes.data, float(cb), tf.ctypes.data, c.ctypes.data) == _lib.OK
    return tf.astype(np.float64), c.astype(np.float64)


def _foci(rng, D):
    low = np.tile(np.array([-2.9671, -1.8326, -2.9671, -3.1416, -2.9671, -0.0873, -2.9671]), D // 7)
    high = -low
    high[3::7], high[5::7] = 0.0873, 3.8223
    return (rng.uniform(low, high).astype(np.float32), rng.uniform(low, high).astype(np.float32))


def test_transform_geometry():
    rng = np.random.default_rng(11)
    for D in (14, 21, 28, 35):
        for trial in range(8):
            f1, f2 = _foci(rng, D)
            cmin = float(np.linalg.norm(f2.astype(np.float64) - f1))
            cb = cmin * float(rng.uniform(1.02, 3.0))
            tf, c = phs(f1, f2, cb)
            r = np.sqrt(cb * cb - cmin * cmin) / 2
            assert np.allclose(c, (f1.astype(np.float64) + f2) / 2, atol=1e-6), "centre is the midpoint of the foci"
            norms = np.linalg.norm(tf, axis=0)
            assert norms[0] == pytest.approx(cb / 2, rel=1e-5), "transverse radius c_best/2"
            assert np.allclose(norms[1:], r, rtol=1e-5), "conjugate radii sqrt(c_best^2 - c_min^2)/2"
            R = tf / norms
            assert np.abs(R.T @ R - np.eye(D)).max() < 1e-5, "rotation is orthogonal"
            axis = (f2.astype(np.float64) - f1) / cmin
            assert np.abs(R[:, 0] - axis).max() < 1e-5, "e_1 maps to the transverse axis f1 -> f2"
            # boundary and interior of the hyperspheroid
            z = rng.normal(size=(256, D))
            z /= np.linalg.norm(z, axis=1, keepdims=True)
            for scale, test in ((1.0, "on"), (0.7, "in")):
                x = c + (z * scale) @ tf.T
                s = np.linalg.norm(x - f1, axis=1) + np.linalg.norm(x - f2, axis=1)
                if test == "on":
                    assert np.allclose(s, cb, rtol=2e-5), "unit sphere maps onto the boundary"
                else:
                    assert (s < cb).all(), "interior maps inside"
            # the foci themselves are inside, the vertices on the boundary
            v = c + tf[:, 0]
            assert np.linalg.norm(v - f1) + np.linalg.norm(v - f2) == pytest.approx(cb, rel=1e-5)


def test_degenerate_cases():
    rng = np.random.default_rng(12)
    f1, _ = _foci(rng, 21)
    tf, c = phs(f1, f1, 3.0)  # coincident foci: a ball of radius c_best/2 about f1
    assert np.allclose(tf, np.eye(21) * 1.5, atol=1e-6)
    assert np.allclose(c, f1, atol=1e-6)
    f1, f2 = _foci(rng, 28)
    cmin = float(np.linalg.norm(f2.astype(np.float64) - f1))
    tf, c = phs(f1, f2, cmin * 0.9)  # best below the focal distance: the segment only
    assert np.allclose(tf[:, 1:], 0.0)
    assert np.linalg.norm(tf[:, 0]) == pytest.approx(cmin * 0.45, rel=1e-5)
    # the transverse axis along e_1 exactly (Householder vector zero): identity rotation
    f1 = np.zeros(14, np.float32)
    f2 = np.zeros(14, np.float32)
    f2[0] = 2.0
    tf, c = phs(f1, f2, 4.0)
    r = np.sqrt(16.0 - 4.0) / 2
    assert np.allclose(tf, np.diag([2.0] + [r] * 13), atol=1e-6)
