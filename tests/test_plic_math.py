"""AL3D / FL3D of the CPU restatement pinned by mathematics alone
(Scardovelli & Zaleski 2000 plane/volume relations), independent of the
reference build:
  - round trip: the volume under the plane AL3D(m, c) is c;
  - symmetry AL3D(m, 1-c) = 1 - AL3D(m, c);
  - axis-aligned normals: alpha = c;
  - FL3D against brute-force midpoint sampling of the cut parallelepiped.
A dropped term, swapped branch or wrong sign in either function fails one
of these."""
import numpy as np

from oracle_ffi import OracleKernels as OK


def test_round_trip_and_symmetry(oracle):
    rng = np.random.default_rng(0)
    for _ in range(3000):
        m = rng.uniform(0, 1, 3)
        if rng.uniform() < 0.2:
            m[rng.integers(3)] = 0.0
        m = m / m.sum()
        c = rng.uniform(0, 1)
        al = OK.al3d(m[0], m[1], m[2], c)
        # volume of {m.x <= al} in the unit cube
        v = OK.fl3d(m[0], m[1], m[2], al, 0.0, 1.0)
        assert abs(v - c) < 1e-11, (m, c, al, v)
        al2 = OK.al3d(m[0], m[1], m[2], 1.0 - c)
        assert abs(al2 - (1.0 - al)) < 1e-11


def test_axis_aligned(oracle):
    for c in (0.0, 1e-9, 0.3, 0.5, 0.77, 1.0 - 1e-9, 1.0):
        for m in ((0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)):
            assert abs(OK.al3d(*m, c) - c) < 1e-15
    # diagonal normal: alpha for c = 1/6 is the corner tetrahedron's (1/3)^(2/3)... alpha = (6*m1*m2*m3*c)^(1/3)
    m = 1.0 / 3.0
    c = 1.0 / 48.0
    assert abs(OK.al3d(m, m, m, c) - (6 * m * m * m * c) ** (1.0 / 3.0)) < 1e-14


def _brute(m, alpha, r0, dr0, n=96):
    # volume of {x: m1 x1 + m2 x2 + m3 x3 <= alpha} for x1 in [r0, r0+dr0], x2,x3 in [0,1]
    t = (np.arange(n) + 0.5) / n
    x1 = r0 + dr0 * t
    X1, X2, X3 = np.meshgrid(x1, t, t, indexing="ij")
    return np.mean(m[0] * X1 + m[1] * X2 + m[2] * X3 <= alpha) * dr0


def test_fl3d_brute_force(oracle):
    rng = np.random.default_rng(1)
    for _ in range(40):
        m = rng.uniform(-1, 1, 3)
        alpha = rng.uniform(-0.5, 1.5)
        r0 = rng.uniform(-0.3, 1.0)
        dr0 = rng.uniform(0.05, 1.0)
        v = OK.fl3d(m[0], m[1], m[2], alpha, r0, dr0)
        vb = _brute(m, alpha, r0, dr0)
        assert abs(v - vb) < 0.02 * dr0 + 1e-3, (m, alpha, r0, dr0, v, vb)
