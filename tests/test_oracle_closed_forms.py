"""Closed-form and invariant checks of the oracle's element kernels, the
project cube root and the rand() restatement: properties the mathematics
fixes, chosen so that a dropped term, a wrong sign or a swapped node fails
one of them."""
import ctypes
import ctypes.util

import numpy as np
import pytest

# reference element node order (images/nodes.png; reference/C/LULESH-OMP.cc:
# 3028-3035): unit-cube corners
UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)


def coords(p):
    return (np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]),
            np.ascontiguousarray(p[:, 2]))


def random_hex(rng, jitter=0.15):
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    if np.linalg.det(A) < 0.2:
        A = np.eye(3)
    p = UNIT @ A.T + rng.standard_normal(3)
    return p + jitter * rng.uniform(-1, 1, p.shape)


def volume(oracle, p):
    x, y, z = coords(p)
    L = oracle.lib()
    return L.lo_elem_volume(oracle.dptr(x), oracle.dptr(y), oracle.dptr(z))


def test_unit_cube_volume(oracle):
    assert volume(oracle, UNIT) == 1.0


def test_affine_map_scales_volume_by_determinant(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        A = rng.standard_normal((3, 3))
        if np.linalg.det(A) <= 0.05:
            continue
        p = UNIT @ A.T + rng.standard_normal(3)
        assert volume(oracle, p) == pytest.approx(np.linalg.det(A), rel=1e-12)


def test_volume_derivative_identities(oracle):
    """sum_i dV/dx_i = 0 (translation) and sum_i x_i . dV/dx_i = 3V (V is
    homogeneous of degree 3), per node against finite differences."""
    rng = np.random.default_rng(2)
    L = oracle.lib()
    for _ in range(20):
        p = random_hex(rng)
        x, y, z = coords(p)
        dx, dy, dz = np.zeros(8), np.zeros(8), np.zeros(8)
        L.lo_elem_volume_derivative(oracle.dptr(dx), oracle.dptr(dy),
                                    oracle.dptr(dz), oracle.dptr(x),
                                    oracle.dptr(y), oracle.dptr(z))
        V = volume(oracle, p)
        scale = np.abs(np.concatenate([dx, dy, dz])).max()
        assert abs(dx.sum()) < 1e-13 * scale
        assert abs(dy.sum()) < 1e-13 * scale
        assert abs(dz.sum()) < 1e-13 * scale
        assert (x @ dx + y @ dy + z @ dz) == pytest.approx(3 * V, rel=1e-12)
        # the volume is trilinear in each node position: central differences
        # are exact up to rounding
        h = 1e-6
        for n in range(8):
            for c, d in ((0, dx), (1, dy), (2, dz)):
                pp, pm = p.copy(), p.copy()
                pp[n, c] += h
                pm[n, c] -= h
                fd = (volume(oracle, pp) - volume(oracle, pm)) / (2 * h)
                assert fd == pytest.approx(d[n], rel=1e-6, abs=1e-9)


def test_node_normals_close_the_surface(oracle):
    """The area-weighted node normals of a closed hexahedron sum to zero, and
    the stress force of a uniform pressure is -p times them."""
    rng = np.random.default_rng(3)
    L = oracle.lib()
    for _ in range(20):
        p = random_hex(rng)
        x, y, z = coords(p)
        nx_, ny_, nz_ = np.zeros(8), np.zeros(8), np.zeros(8)
        L.lo_elem_node_normals(oracle.dptr(nx_), oracle.dptr(ny_),
                               oracle.dptr(nz_), oracle.dptr(x),
                               oracle.dptr(y), oracle.dptr(z))
        for v in (nx_, ny_, nz_):
            assert abs(v.sum()) < 1e-13 * np.abs(v).max()
    # divergence theorem, sum_i x_i n_x,i = V: exact when every face is a
    # parallelogram (vertex average = face centroid), i.e. affine elements
    for _ in range(20):
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(A) <= 0.1:
            continue
        p = UNIT @ A.T + rng.standard_normal(3)
        x, y, z = coords(p)
        nx_, ny_, nz_ = np.zeros(8), np.zeros(8), np.zeros(8)
        L.lo_elem_node_normals(oracle.dptr(nx_), oracle.dptr(ny_),
                               oracle.dptr(nz_), oracle.dptr(x),
                               oracle.dptr(y), oracle.dptr(z))
        V = np.linalg.det(A)
        assert x @ nx_ == pytest.approx(V, rel=1e-12)
        assert y @ ny_ == pytest.approx(V, rel=1e-12)
        assert z @ nz_ == pytest.approx(V, rel=1e-12)
        assert x @ ny_ == pytest.approx(0.0, abs=1e-12)


def test_jacobian_volume_of_parallelepiped(oracle):
    """For an affine element the shape-function Jacobian determinant equals
    the exact volume."""
    rng = np.random.default_rng(4)
    L = oracle.lib()
    for _ in range(20):
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(A) <= 0.1:
            continue
        p = UNIT @ A.T
        x, y, z = coords(p)
        b = np.zeros(24)
        vol = L.lo_elem_shape_derivatives(oracle.dptr(x), oracle.dptr(y),
                                          oracle.dptr(z), oracle.dptr(b))
        assert vol == pytest.approx(np.linalg.det(A), rel=1e-12)


def test_hourglass_force_vanishes_for_affine_velocity(oracle):
    """The Flanagan-Belytschko modes are orthogonal to linear velocity
    fields; a non-linear (hourglass) field is resisted."""
    rng = np.random.default_rng(5)
    L = oracle.lib()
    for _ in range(20):
        p = random_hex(rng, jitter=0.1)
        x, y, z = coords(p)
        G = rng.standard_normal((3, 3))
        v = p @ G.T + rng.standard_normal(3)
        xd, yd, zd = coords(v)
        fx, fy, fz = np.zeros(8), np.zeros(8), np.zeros(8)
        V = volume(oracle, p)
        L.lo_elem_hourglass_force(*(oracle.dptr(a) for a in (x, y, z, xd, yd, zd)),
                                  V, -1.0, *(oracle.dptr(a) for a in (fx, fy, fz)))
        scale = np.abs(v).max()
        assert np.abs(np.concatenate([fx, fy, fz])).max() < 1e-11 * scale
        # the pure hourglass mode Gamma_0 on the x velocity
        g0 = np.array([1, 1, -1, -1, -1, -1, 1, 1], np.float64)
        xd2 = np.ascontiguousarray(g0)
        zero = np.zeros(8)
        L.lo_elem_hourglass_force(*(oracle.dptr(a) for a in (x, y, z, xd2, zero, zero)),
                                  V, -1.0, *(oracle.dptr(a) for a in (fx, fy, fz)))
        assert fx @ xd2 < 0  # dissipative: the force opposes the mode


def test_project_cbrt_accuracy(oracle):
    """The project cube root is within half an ulp (plus the rounding of the
    check) of the exact root, and agrees with the C library's to 1 ulp."""
    rng = np.random.default_rng(6)
    L = oracle.lib()
    xs = np.concatenate([
        np.ldexp(1 + rng.uniform(0, 1, 20000), rng.integers(-1000, 1000, 20000)),
        np.array([1.0, 8.0, 27.0, 1e-300, 5e-324, 1.7e308, 0.125, 3.375e-3]),
    ])
    for x in xs:
        r = L.lo_cbrt_project(float(x))
        c = np.cbrt(x)
        assert abs(r - c) <= np.spacing(c) * 1.0 + 0.0, (x, r, c)
    # exact cubes
    for k in range(1, 2000):
        assert L.lo_cbrt_project(float(k) ** 3) == float(k)
    assert L.lo_cbrt_project(-8.0) == -2.0
    assert L.lo_cbrt_project(0.0) == 0.0


def test_project_cbrt_is_correctly_rounded(oracle):
    """r is the binary64 value nearest to x^(1/3) exactly when x lies strictly
    between the cubes of the midpoints to r's neighbours; checked in rational
    arithmetic over the whole exponent range (subnormals included) and at the
    ends of the reduction interval."""
    from fractions import Fraction
    rng = np.random.default_rng(11)
    L = oracle.lib()
    xs = list(np.ldexp(1 + rng.uniform(0, 1, 4000), rng.integers(-1074, 1024, 4000)))
    xs += [np.nextafter(8.0, 0.0), np.nextafter(1.0, 0.0), np.nextafter(1.0, 2.0),
           2.0, 4.0, 5e-324, np.nextafter(2.2250738585072014e-308, 0.0),
           2.2250738585072014e-308, 1.7976931348623157e308]
    for x in xs:
        x = float(x)
        if not (0.0 < x < np.inf):
            continue
        r = L.lo_cbrt_project(x)
        below = Fraction(r - np.nextafter(r, 0.0)) / 2
        above = Fraction(np.nextafter(r, np.inf) - r) / 2
        assert (Fraction(r) - below) ** 3 < Fraction(x) < (Fraction(r) + above) ** 3, x.hex()
        assert L.lo_cbrt_project(-x) == -r


def test_rand_restatement_matches_c_library(oracle):
    """The region layout draws from rand() after srand(0); the restatement
    reproduces the C library's sequence (glibc here and on the GPU host)."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    for seed in (0, 1, 12345):
        libc.srand(seed)
        want = np.array([libc.rand() for _ in range(5000)], np.int32)
        np.testing.assert_array_equal(oracle.rand_sequence(seed, 5000), want)
