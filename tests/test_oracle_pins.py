"""Pins of the CPU oracle against what the mathematics and the reference's own
code fix, independent of the oracle's own formulas:

* Go conversion semantics: special values with known answers.
* 4x4 transform: exact-integer closed forms, an association-order known answer,
  and the reference's c/matrix_amd64.c (compiled into oracle/_ref) bit-for-bit
  for the pairwise variant, plus its c/main.c known answer "4 3 2 1".
* Matrix builders: closed forms with exactly representable results.
* Clipping: identity / rejection closed forms, geometric invariants and an
  exact-rational (fractions.Fraction) Sutherland-Hodgman.
* Raster: a split rectangle covers each pixel exactly once (top-left rule),
  brute-force coverage of tiny triangles, exact barycentrics, u/v sign.
* Sampling and colour scaling: exhaustive grids.
"""
import ctypes
import math
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

from oracle import refcpu
from oracle.refcpu import _p

F32 = np.float32
INT64_MIN = -(1 << 63)


@pytest.fixture(scope="module")
def lib():
    return refcpu.load()


def f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------------------
# Go conversion semantics
# ---------------------------------------------------------------------------

def test_go_int(lib):
    cases = {1.9: 1, -1.9: -1, -0.5: 0, 0.0: 0, 1e10: 10000000000, float("nan"): INT64_MIN,
             float("inf"): INT64_MIN, float("-inf"): INT64_MIN, 2.0 ** 63: INT64_MIN,
             -(2.0 ** 63): INT64_MIN, 2.0 ** 63 - 2.0 ** 39: (1 << 63) - (1 << 39),
             -3.0e9: -3000000000}
    for x, want in cases.items():
        assert lib.rc_go_int(x) == want, x


def test_go_uint8(lib):
    cases = {255.9: 255, 256.0: 0, 257.5: 1, -1.0: 255, -0.5: 0, -2.5: 254,
             float("nan"): 0, 2.0 ** 31: 0, -(2.0 ** 31): 0, 3e9: 0, 65791.0: 255,
             float("inf"): 0}
    for x, want in cases.items():
        assert lib.rc_go_uint8(x) == want, x


def test_go_min_max(lib):
    nan = float("nan")
    for a, b in [(nan, 1.0), (1.0, nan)]:
        assert math.isnan(lib.rc_go_min(a, b)) and math.isnan(lib.rc_go_max(a, b))
    assert math.copysign(1, lib.rc_go_min(0.0, -0.0)) < 0
    assert math.copysign(1, lib.rc_go_min(-0.0, 0.0)) < 0
    assert math.copysign(1, lib.rc_go_max(-0.0, 0.0)) > 0
    assert math.copysign(1, lib.rc_go_max(0.0, -0.0)) > 0
    assert lib.rc_go_min(2.0, -3.0) == -3.0 and lib.rc_go_max(2.0, -3.0) == 2.0


def test_sqrt32_correctly_rounded(lib):
    rng = np.random.default_rng(0)
    xs = rng.random(2000).astype(np.float32) * F32(1e6)
    for x in xs[:500]:
        r = lib.rc_sqrt32(float(x))
        # correctly rounded: r^2 brackets x within half an ulp of r (exact rationals)
        lo = Fraction(float(np.nextafter(F32(r), F32(0)))) + Fraction(r)
        hi = Fraction(float(np.nextafter(F32(r), F32(np.inf)))) + Fraction(r)
        assert (lo / 2) ** 2 <= Fraction(float(x)) <= (hi / 2) ** 2


# ---------------------------------------------------------------------------
# Transform
# ---------------------------------------------------------------------------

def transform(lib, m, vecs, pairwise=False):
    mm = refcpu.matrix(m)
    v = np.ascontiguousarray(vecs, np.float32).copy()
    fn = lib.rc_transform_batch_pairwise if pairwise else lib.rc_transform_batch
    fn(ctypes.byref(mm), _p(v), len(v))
    return v


def test_transform_exact_integers(lib):
    """Small integers: every product and sum is exact, so the result is the
    integer matrix-vector product whatever the association."""
    rng = np.random.default_rng(1)
    m = rng.integers(-50, 50, (4, 4))
    vs = rng.integers(-50, 50, (64, 4))
    want = vs @ m.T  # row i of the output = M @ v
    for pw in (False, True):
        got = transform(lib, m.astype(np.float32), vs.astype(np.float32), pw)
        assert np.array_equal(got, want.astype(np.float32))


def test_transform_association_order(lib):
    """asm_purego.go:9-19 sums left to right: ((2^24 + 0) + 1) + 1 = 2^24 in
    float32 (each +1 ties to even), while the SSE form pairs
    (2^24 + 0) + (1 + 1) = 2^24 + 2."""
    m = np.zeros((4, 4), np.float32)
    m[0] = [2.0 ** 24, 0, 1, 1]
    v = np.ones((1, 4), np.float32)
    assert transform(lib, m, v)[0, 0] == 2.0 ** 24
    assert transform(lib, m, v, pairwise=True)[0, 0] == 2.0 ** 24 + 2


def _aligned(n_floats):
    buf = np.zeros(n_floats + 16, np.float32)
    off = (-buf.ctypes.data % 16) // 4
    return buf[off:off + n_floats]


@pytest.mark.skipif(not os.path.exists(refcpu.REF_LIB_PATH),
                    reason="oracle/_ref not built (needs the reference tree)")
def test_pairwise_matches_reference_c(lib):
    """The reference's own c/matrix_amd64.c, compiled unmodified, agrees bit for
    bit with the oracle's pairwise variant (mat4 columns = transposed rows,
    asm_amd64.go:9)."""
    ref = ctypes.CDLL(refcpu.REF_LIB_PATH)
    ref.matrix_multiply_vec4.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    ref.matrix_multiply_vec4.restype = None
    rng = np.random.default_rng(2)
    m = (rng.standard_normal((4, 4)) * 3).astype(np.float32)
    vs = (rng.standard_normal((1000, 4)) * 100).astype(np.float32)
    cols = _aligned(16)
    cols[:] = m.T.reshape(16)
    out = vs.copy()
    ref.matrix_multiply_vec4(cols.ctypes.data, out.ctypes.data, len(out))
    assert np.array_equal(out.view(np.uint32), transform(lib, m, vs, True).view(np.uint32))

    # c/main.c: rows _mm_set_ps(1,0,0,0) .. _mm_set_ps(0,0,0,4) times (1,1,1,1)
    cols[:] = np.array([[0, 0, 0, 1], [0, 0, 2, 0], [0, 3, 0, 0], [4, 0, 0, 0]],
                       np.float32).reshape(16)
    one = np.ones((1, 4), np.float32)
    ref.matrix_multiply_vec4(cols.ctypes.data, one.ctypes.data, 1)
    assert one.tolist() == [[4.0, 3.0, 2.0, 1.0]]


def test_matrix_multiply_closed_form(lib):
    rng = np.random.default_rng(3)
    a = rng.integers(-9, 9, (4, 4))
    b = rng.integers(-9, 9, (4, 4))
    ma, mb = refcpu.matrix(a), refcpu.matrix(b)
    r = refcpu.matrix_np(lib.rc_matrix_multiply(ctypes.byref(ma), ctypes.byref(mb)))
    assert np.array_equal(r, (a @ b).astype(np.float32))
    # res starts at +0 (matrix.go:155-165): (-0)*(1) summed onto +0 is +0
    z = refcpu.matrix(np.full((4, 4), -0.0, np.float32))
    o = refcpu.matrix(np.ones((4, 4), np.float32))
    r = refcpu.matrix_np(lib.rc_matrix_multiply(ctypes.byref(z), ctypes.byref(o)))
    assert not np.signbit(r).any()


def test_builders_closed_forms(lib):
    V3 = refcpu.rc_vec3
    # angle 0 short-circuits to identity (matrix.go:33-35)
    r = refcpu.matrix_np(lib.rc_rotation_matrix(0.0, 0.0, 0.0))
    assert np.array_equal(r, np.eye(4, dtype=np.float32))
    w = refcpu.matrix_np(lib.rc_world_matrix(V3(2, 3, 4), V3(0, 0, 0), V3(5, 6, 7)))
    want = np.array([[2, 0, 0, 5], [0, 3, 0, 6], [0, 0, 4, 7], [0, 0, 0, 1]], np.float32)
    assert np.array_equal(w, want)
    s = refcpu.matrix_np(lib.rc_screen_matrix(800, 600))
    assert np.array_equal(s, np.array([[400, 0, 0, 400], [0, 300, 0, 300], [0, 0, .5, .5],
                                       [0, 0, 0, 1]], np.float32))
    # znear 1, zfar 3: m22 = 4/-2 = -2, m23 = 6/-2 = -3, stored negated (matrix.go:97-104)
    p = refcpu.matrix_np(lib.rc_perspective_matrix(f32(math.pi / 2), 1.0, 1.0, 3.0))
    assert p[2, 2] == 2 and p[2, 3] == 3 and p[3, 2] == -1 and p[3, 3] == 0
    assert abs(p[0, 0] - 1) < 1e-6 and abs(p[1, 1] - 1) < 1e-6
    # view: eye (1,2,3) looking down -z with +y up: x axis (-1,0,0), y (0,1,0), z (0,0,-1)
    v = refcpu.matrix_np(lib.rc_view_matrix(V3(1, 2, 3), V3(0, 0, -1), V3(0, 1, 0)))
    assert np.array_equal(v, np.array([[-1, 0, 0, 1], [0, 1, 0, -2], [0, 0, -1, 3],
                                       [0, 0, 0, 1]], np.float32))
    # rotation by pi/2 about z maps x to y up to float rounding of sin/cos
    r = refcpu.matrix_np(lib.rc_rotation_matrix(0.0, 0.0, f32(math.pi / 2)))
    assert np.allclose(r[:2, :2], [[0, -1], [1, 0]], atol=1e-7)


def test_face_normals_and_bbox(lib):
    v = np.array([[0, 0, 0, 1], [2, 0, 0, 1], [0, 2, 0, 1], [0, 0, 2, 1]], np.float32)
    f = np.array([[0, 1, 2], [0, 3, 1]], np.int32)
    n = refcpu.face_normals(v, f)
    assert n.tolist() == [[0, 0, 1, 1], [0, 1, 0, 1]]
    b = refcpu.bounding_box(np.array([[1, -2, 3, 1], [-1, 5, 0.5, 7], [0, 0, -0.0, 1]],
                                     np.float32))
    xs, ys, zs = sorted(set(b[:, 0])), sorted(set(b[:, 1])), sorted(set(b[:, 2]))
    assert xs == [-1, 1] and ys == [-2, 5] and zs == [0, 3]
    assert (b[:, 3] == 1).all() and len({tuple(r) for r in b}) == 8
    # corner order: x outer, then y, then z (mesh.go:40-49)
    assert b[0].tolist()[:3] == [-1, -2, 0] and b[1].tolist()[:3] == [-1, -2, 3]
    assert b[2].tolist()[:3] == [-1, 5, 0] and b[4].tolist()[:3] == [1, -2, 0]
    nanb = refcpu.bounding_box(np.array([[1, 1, 1, 1], [np.nan, 0, 0, 1]], np.float32))
    assert np.isnan(nanb[:, 0]).all()


# ---------------------------------------------------------------------------
# Frustum
# ---------------------------------------------------------------------------

def box(xs, ys, zs, w=-1.0):
    return np.array([[x, y, z, w] for x in xs for y in ys for z in zs], np.float32)


def test_box_visibility_hand_classified(lib):
    INSIDE, INTERSECT, OUTSIDE = 2, 1, 0
    # with W = -1 the inside region is |x|,|y| <= 1, znear <= z <= zfar
    assert lib.rc_box_visibility(0.0, 50.0, _p(box((-.5, .5), (-.5, .5), (1, 2)))) == INSIDE
    assert lib.rc_box_visibility(0.0, 50.0, _p(box((2, 3), (-.5, .5), (1, 2)))) == OUTSIDE
    assert lib.rc_box_visibility(0.0, 50.0, _p(box((-2, .5), (-.5, .5), (1, 2)))) == INTERSECT
    assert lib.rc_box_visibility(0.0, 50.0, _p(box((-.5, .5), (-.5, .5), (-2, -1)))) == OUTSIDE
    assert lib.rc_box_visibility(0.0, 1.5, _p(box((-.5, .5), (-.5, .5), (1, 2)))) == INTERSECT
    # the first plane that sees a corner outside decides (clipping.go:147-149):
    # straddling Left/Right but wholly above Top is still Intersect
    assert lib.rc_box_visibility(0.0, 50.0, _p(box((-2, .5), (2, 3), (1, 2)))) == INTERSECT
    # W = +1 boxes can never be inside (x <= -w and x >= w)
    assert lib.rc_box_visibility(0.0, 50.0, _p(box((-.5, .5), (-.5, .5), (1, 2), 1.0))) != INSIDE


def test_facing_camera(lib):
    a = np.array([[0, 0, 1, 1], [1, 0, 1, 1], [0, 1, 1, 1]], np.float32)
    assert lib.rc_facing_camera(_p(a)) == 0          # normal +z, v0 at z=+1: dot = -1
    assert lib.rc_facing_camera(_p(a[[0, 2, 1]].copy())) == 1
    flat = np.array([[0, 0, 1, 1], [1, 0, 1, 1], [2, 0, 1, 1]], np.float32)
    assert lib.rc_facing_camera(_p(flat)) == 0       # strictly > 0


def clip(lib, pts, uvs, ints, zn=0.0, zf=50.0):
    p = np.ascontiguousarray(pts, np.float32)
    u = np.ascontiguousarray(uvs, np.float32)
    i = np.ascontiguousarray(ints, np.float32)
    op = np.zeros((9, 3, 4), np.float32)
    ou = np.zeros((9, 3, 2), np.float32)
    oi = np.zeros((9, 3), np.float32)
    n = lib.rc_clip_triangle(zn, zf, _p(p), _p(u), _p(i), _p(op), _p(ou), _p(oi))
    return n, op[:max(n, 0)], ou[:max(n, 0)], oi[:max(n, 0)]


PLANES = [((-1, 0, 0, 1), (1, 0, 0, 1)), ((1, 0, 0, 1), (-1, 0, 0, 1)),
          ((0, -1, 0, 1), (0, 1, 0, 1)), ((0, 1, 0, 1), (0, -1, 0, 1))]


def exact_clip(pts, uvs, ints, zn, zf):
    """Sutherland-Hodgman on exact rationals: same planes, edge order and fan."""
    planes = PLANES + [((0, 0, zn, 1), (0, 0, -1, 0)), ((0, 0, zf, 1), (0, 0, 1, 0))]
    planes = [([Fraction(c) for c in P], [Fraction(c) for c in N]) for P, N in planes]
    poly = [([Fraction(float(c)) for c in pts[k]] + [Fraction(float(c)) for c in uvs[k]]
             + [Fraction(float(ints[k]))]) for k in range(3)]

    def inside(P, N, q):
        return sum((q[i] - P[i]) * N[i] for i in range(4)) <= 0

    for P, N in planes:
        out = []
        n = len(poly)
        for b in range(n):
            A, B = poly[(b + 1) % n], poly[b]
            ia, ib = inside(P, N, A), inside(P, N, B)
            if ia != ib:
                d = sum(N[i] * (B[i] - A[i]) for i in range(4))
                t = -sum(N[i] * (A[i] - P[i]) for i in range(4)) / d
                X = [A[i] + (B[i] - A[i]) * t for i in range(7)]
            if ia:
                if not ib:
                    out.append(X)
                out.append(A)
            elif ib:
                out.append(X)
        if not out:
            return []
        poly = out
    return [(poly[0], poly[i + 1], poly[i + 2]) for i in range(len(poly) - 2)]


def test_clip_identity_and_rejection(lib):
    tri = np.array([[0.1, 0.2, 1, -2], [-0.5, 0.3, 2, -2], [0.2, -0.4, 3, -2]], np.float32)
    uv = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    it = np.array([0.1, 0.2, 0.3], np.float32)
    n, op, ou, oi = clip(lib, tri, uv, it)
    assert n == 1 and np.array_equal(op[0], tri) and np.array_equal(ou[0], uv)
    assert np.array_equal(oi[0], it)
    far = tri.copy()
    far[:, 0] += 10
    assert clip(lib, far, uv, it)[0] == 0
    behind = tri.copy()
    behind[:, 2] = -1
    assert clip(lib, behind, uv, it)[0] == 0


def test_clip_closed_form_left_plane(lib):
    """Triangle with W=-1 crossing x = 1 (the 'Left' plane x <= -w) at exactly
    representable points: (0,0)-(2,0)-(0,2) clipped to x <= 1."""
    tri = np.array([[0, 0, 1, -1], [2, 0, 1, -1], [0, 0.5, 1, -1]], np.float32)
    uv = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    it = np.array([0.0, 1.0, 0.5], np.float32)
    n, op, ou, oi = clip(lib, tri, uv, it)
    pts = {tuple(p[:2]) for t in op for p in t}
    assert pts == {(0, 0), (1, 0), (1, 0.25), (0, 0.5)} and n == 2
    for t, u, i in zip(op, ou, oi):
        for k in range(3):
            if tuple(t[k][:2]) == (1.0, 0.0):
                assert u[k].tolist() == [0.5, 0] and i[k] == 0.5


def test_clip_random_vs_exact(lib):
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(400):
        pts = np.concatenate([rng.uniform(-3, 3, (3, 3)), rng.uniform(-3, 0.5, (3, 1))], 1)
        pts[:, 2] = rng.uniform(-1, 5, 3)
        pts = pts.astype(np.float32)
        uv = rng.uniform(-1, 2, (3, 2)).astype(np.float32)
        it = rng.uniform(0, 1, 3).astype(np.float32)
        n, op, ou, oi = clip(lib, pts, uv, it, 0.0, 4.0)
        # skip near-degenerate inputs where float and exact classification may differ
        dist = []
        for P, N in PLANES + [((0, 0, 0, 1), (0, 0, -1, 0)), ((0, 0, 4, 1), (0, 0, 1, 0))]:
            for q in pts:
                dist.append(abs(sum((float(q[i]) - P[i]) * N[i] for i in range(4))))
        if min(dist) < 1e-3:
            continue
        want = exact_clip(pts, uv, it, 0, 4)
        assert n == len(want)
        assert 0 <= n <= 7
        for t, tu, ti, w in zip(op, ou, oi, want):
            for k in range(3):
                got = list(t[k]) + list(tu[k]) + [ti[k]]
                for g, e in zip(got, w[k]):
                    assert abs(g - float(e)) <= 2e-4 * max(1.0, abs(float(e))), (got, w[k])
        checked += 1
    assert checked > 150


def test_clip_output_attributes_are_affine(lib):
    """Every clipped vertex carries the uv/intensity the input triangle's affine
    interpolation assigns to its position (catches swapped lerp endpoints)."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        pts = np.concatenate([rng.uniform(-2, 2, (3, 2)), rng.uniform(0.5, 3, (3, 1)),
                              np.full((3, 1), -1.0)], 1).astype(np.float32)
        uv = rng.uniform(0, 1, (3, 2)).astype(np.float32)
        it = rng.uniform(0, 1, 3).astype(np.float32)
        n, op, ou, oi = clip(lib, pts, uv, it)
        A = np.stack([pts[:, 0], pts[:, 1], np.ones(3)]).astype(np.float64)
        if abs(np.linalg.det(A)) < 1e-3:
            continue
        for t, tu, ti in zip(op, ou, oi):
            for k in range(3):
                bary = np.linalg.solve(A, [t[k][0], t[k][1], 1.0])
                assert np.allclose(bary @ uv, tu[k], atol=1e-4)
                assert abs(bary @ it - ti[k]) < 1e-4
                # and lies inside all planes (|x|,|y| <= |w|), within rounding
                assert abs(t[k][0]) <= 1 + 1e-5 and abs(t[k][1]) <= 1 + 1e-5


# ---------------------------------------------------------------------------
# Raster
# ---------------------------------------------------------------------------

def raster(lib, w, h, verts, zs=(-1, -1, -1), uvs=((0, 0),) * 3, ints=(1, 1, 1), tex=None,
           color=(200, 200, 200, 255), rgba=None, depth=None):
    if rgba is None:
        rgba = np.zeros((h, w, 4), np.uint8)
        depth = np.full((h, w), -1.0, np.float32)
    fc = np.array(color, np.uint8)
    passed = ctypes.c_int64()
    (x0, y0), (x1, y1), (x2, y2) = verts
    cov = lib.rc_triangle(w, h, _p(rgba), _p(depth), x0, y0, zs[0], *uvs[0], x1, y1, zs[1],
                          *uvs[1], x2, y2, zs[2], *uvs[2], 0, 0, w, h, *ints,
                          None if tex is None else ctypes.byref(tex), _p(fc),
                          ctypes.byref(passed), None)
    return cov, rgba, depth


def coverage_mask(lib, w, h, verts):
    _, rgba, _ = raster(lib, w, h, verts)
    return rgba[..., 3] == 255


def test_split_rectangle_covers_each_pixel_once(lib):
    """Two triangles sharing a diagonal tile an axis-aligned rectangle: the
    top-left rule gives every lattice point exactly one owner."""
    rng = np.random.default_rng(6)
    for _ in range(30):
        x0, y0 = rng.integers(0, 10, 2)
        x1, y1 = x0 + rng.integers(1, 12), y0 + rng.integers(1, 12)
        # one winding only (f < 0, rasterizer.go:147): pick the one that fills
        quads = [((x0, y0), (x1, y1), (x1, y0)), ((x0, y0), (x0, y1), (x1, y1))]
        cnt = np.zeros((32, 32), int)
        for t in quads:
            cnt += coverage_mask(lib, 32, 32, t)
        assert cnt.max() == 1
        assert cnt.sum() == (x1 - x0) * (y1 - y0)


def test_shared_edges_fan_no_gaps(lib):
    """A fan of triangles around a centre tiles its hull without overlap."""
    c = (20, 20)
    ring = [(35, 20), (30, 32), (14, 36), (5, 25), (9, 8), (24, 4)]
    cnt = np.zeros((48, 48), int)
    for i in range(len(ring)):
        a, b = ring[i], ring[(i + 1) % len(ring)]
        for t in ((c, b, a), (c, a, b)):
            m = coverage_mask(lib, 48, 48, t)
            if m.any():
                cnt += m
                break
    assert cnt.max() == 1


def test_coverage_brute_force(lib):
    """Coverage equals the direct per-pixel edge-function test of
    rasterizer.go:107-147 evaluated with Python integers."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        v = [tuple(int(c) for c in rng.integers(-3, 20, 2)) for _ in range(3)]
        got = coverage_mask(lib, 16, 16, v)
        want = np.zeros((16, 16), bool)
        (x0, y0), (x1, y1), (x2, y2) = v
        edges = [(y0 - y1, x1 - x0, x0 * y1 - x1 * y0), (y1 - y2, x2 - x1, x1 * y2 - x2 * y1),
                 (y2 - y0, x0 - x2, x2 * y0 - x0 * y2)]
        for y in range(max(min(y0, y1, y2), 0), min(max(y0, y1, y2), 15) + 1):
            for x in range(max(min(x0, x1, x2), 0), min(max(x0, x1, x2), 15) + 1):
                ok = True
                for A, B, C in edges:
                    f = A * x + B * y + C
                    if not (B > 0 or (B == 0 and A > 0)):
                        f -= 1
                    ok = ok and f < 0
                want[y, x] = ok
        assert np.array_equal(got, want), v


def test_barycentrics_and_depth_exact(lib):
    """Right triangle (0,0),(0,8),(8,0) wound for f<0: alpha = f12/S,
    beta = f20/S with the adjusted edge values; with z = -2,-4,-8 the depth is
    -(a/z0 + b/z1 + g/z2) rounded step by step."""
    verts = [(0, 0), (0, 8), (8, 0)]
    zs = (-2.0, -4.0, -8.0)
    _, _, depth = raster(lib, 16, 16, verts, zs=zs)
    (x0, y0), (x1, y1), (x2, y2) = verts
    e01 = (y0 - y1, x1 - x0, x0 * y1 - x1 * y0)
    e12 = (y1 - y2, x2 - x1, x1 * y2 - x2 * y1)
    e20 = (y2 - y0, x0 - x2, x2 * y0 - x0 * y2)

    def f(e, x, y):
        A, B, C = e
        v = A * x + B * y + C
        return v if (B > 0 or (B == 0 and A > 0)) else v - 1

    n = 0
    for y in range(16):
        for x in range(16):
            a, b, c = f(e01, x, y), f(e12, x, y), f(e20, x, y)
            if a < 0 and b < 0 and c < 0:
                S = F32(b + c + a)
                al, be = F32(b) / S, F32(c) / S
                ga = (F32(1) - al) - be
                z = -((al / F32(zs[0]) + be / F32(zs[1])) + ga / F32(zs[2]))
                assert depth[y, x] == z
                n += 1
            else:
                assert depth[y, x] == -1
    assert n > 20


def test_uv_sign_and_flip(lib):
    """u,v come out negated (rasterizer.go:153,158-159) and sampling flips u
    with (1-u) (texture.go:75): a constant u = v = 0.3 at z = -1 reads column
    int(1.3*8) & 7 = 2 (not int(0.7*8) = 5) and row int(-0.3*8) & 7 = 6 (not 2)."""
    px = np.zeros((8, 8, 4), np.uint8)
    px[..., 0] = 30 * np.arange(8)[None, :] + 15   # R encodes the column
    px[..., 1] = 30 * np.arange(8)[:, None] + 15   # G encodes the row
    px[..., 3] = 255
    t = refcpu.rc_texture(2, 8, 8, 8.0, 8.0, 1.0, (ctypes.c_uint8 * 4)(), px.ctypes.data)
    verts = [(0, 0), (0, 12), (12, 0)]
    _, rgba, _ = raster(lib, 16, 16, verts, uvs=((0.3, 0.3),) * 3, tex=t)
    inside = rgba[..., 3] == 255
    assert inside.any()
    # intensity alpha+beta+gamma may round to 1 - 2^-24, so decode with // 30
    assert set(rgba[inside][:, 0] // 30) == {2}
    assert set(rgba[inside][:, 1] // 30) == {6}


def test_depth_test_ties_last_wins(lib):
    """zRec >= Z: an equal-depth later fragment overwrites (rasterizer.go:156)."""
    rgba = np.zeros((8, 8, 4), np.uint8)
    depth = np.full((8, 8), -1.0, np.float32)
    v = [(0, 0), (0, 8), (8, 0)]
    raster(lib, 8, 8, v, color=(10, 0, 0, 255), rgba=rgba, depth=depth)
    raster(lib, 8, 8, v, color=(20, 0, 0, 255), rgba=rgba, depth=depth)
    m = rgba[..., 3] == 255
    assert m.any() and set(rgba[m][:, 0]) == {20}
    raster(lib, 8, 8, v, zs=(-1.5, -1.5, -1.5), color=(30, 0, 0, 255), rgba=rgba, depth=depth)
    assert set(rgba[m][:, 0]) == {20}              # farther (smaller 1/z) loses


# ---------------------------------------------------------------------------
# Sampling / colour
# ---------------------------------------------------------------------------

def trunc_i64(x):
    x = float(x)
    if not (-(2.0 ** 63) <= x < 2.0 ** 63):
        return INT64_MIN
    return int(x)


def test_sample_grid(lib):
    rng = np.random.default_rng(8)
    for (w, h, kind, scale) in [(8, 4, 2, 1.0), (8, 8, 2, 2.5), (5, 3, 1, 1.0), (7, 6, 1, 0.3)]:
        px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        t = refcpu.rc_texture(kind, w, h, float(w), float(h), scale, (ctypes.c_uint8 * 4)(),
                              px.ctypes.data)
        out = np.zeros(4, np.uint8)
        for u in np.linspace(-3.3, 3.3, 37, dtype=np.float32):
            for v in np.linspace(-2.2, 2.7, 29, dtype=np.float32):
                lib.rc_sample(ctypes.byref(t), float(u), float(v), _p(out))
                xf = (F32(1) - u) * F32(scale) * F32(w)
                yf = v * F32(scale) * F32(h)
                if kind == 2:
                    x, y = trunc_i64(xf) & (w - 1), trunc_i64(yf) & (h - 1)
                    idx = y * w + x
                else:
                    x = math.fmod(trunc_i64(xf), w)
                    y = math.fmod(trunc_i64(yf), h)
                    idx = int(y) * w + int(x)
                    idx = max(idx, 0)
                assert out.tolist() == px.reshape(-1, 4)[idx].tolist(), (w, h, u, v)
    solid = refcpu.rc_texture(0, 0, 0, 0, 0, 1.0, (ctypes.c_uint8 * 4)(1, 2, 3, 4), None)
    lib.rc_sample(ctypes.byref(solid), 0.3, 0.7, _p(out))
    assert out.tolist() == [1, 2, 3, 4]
    weird = refcpu.rc_texture(7, 0, 0, 0, 0, 1.0, (ctypes.c_uint8 * 4)(), None)
    lib.rc_sample(ctypes.byref(weird), 0.3, 0.7, _p(out))
    assert out.tolist() == [255, 0, 255, 255]


def test_color_intensity_exhaustive(lib):
    out = np.zeros(4, np.uint8)
    ks = np.array([0, 0.25, 0.5, 0.999, 1.0, 1.0039216, 1.01, 2.0, -0.001, -0.5, -1.2,
                   float("nan")], np.float32)
    for c in range(256):
        col = np.array([c, 255 - c, c // 2, 77], np.uint8)
        for k in ks:
            lib.rc_color_intensity(_p(col), float(k), _p(out))
            for ch in range(3):
                p = F32(col[ch]) * k
                if np.isnan(p) or not (-(2.0 ** 31) <= float(p) < 2.0 ** 31):
                    want = 0
                else:
                    want = int(float(p)) & 0xFF
                assert out[ch] == want
            assert out[3] == 77


def test_clear_and_dot_grid(lib):
    w, h = 37, 23
    rgba = np.zeros((h, w, 4), np.uint8)
    depth = np.zeros((h, w), np.float32)
    lib.rc_clear(w, h, _p(rgba), _p(depth), _p(np.array([50, 50, 50, 255], np.uint8)))
    lib.rc_dot_grid(w, h, _p(rgba), _p(np.array([100, 100, 100, 255], np.uint8)), 10)
    assert (depth == -1).all()
    dots = {(y, x) for y in (10, 20) for x in (10, 20, 30)}
    for y in range(h):
        for x in range(w):
            want = 100 if (y, x) in dots else 50
            assert rgba[y, x, 0] == want


def test_tile_bounds(lib):
    """calculateTileBoundaries (renderer.go:50-76): 4x4 grid of ceil sizes,
    inclusive ends clamped to the framebuffer."""
    s, e = refcpu.rc_vec2(), refcpu.rc_vec2()
    lib.rc_tile_bounds(0, 1, 1280, 720, ctypes.byref(s), ctypes.byref(e))
    assert (s.x, s.y, e.x, e.y) == (0, 0, 1280, 720)
    lib.rc_tile_bounds(5, 16, 801, 601, ctypes.byref(s), ctypes.byref(e))
    assert (s.x, s.y, e.x, e.y) == (201, 151, 402, 302)
    lib.rc_tile_bounds(15, 16, 801, 601, ctypes.byref(s), ctypes.byref(e))
    assert (s.x, s.y, e.x, e.y) == (603, 453, 801, 601)


# ---- overlays: FrameBuffer.Pixel / Rect / Line (rasterizer.go:25-30,54-79), abs (utils.go) ----

def _overlay_buf(w, h):
    return np.zeros((h, w, 4), np.uint8)


def _written(buf):
    return {(int(x), int(y)) for y, x in zip(*np.nonzero(buf[..., 3]))}


def test_go_abs_and_pixel_wrap():
    lib = refcpu.load()
    assert lib.rc_go_abs(-5) == 5 and lib.rc_go_abs(7) == 7
    assert lib.rc_go_abs(-(1 << 63)) == (1 << 63) - 1  # minInt -> maxInt
    w, h = 5, 4
    c = (ctypes.c_uint8 * 4)(1, 2, 3, 255)
    buf = _overlay_buf(w, h)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    lib.rc_pixel(w, h, p, 0, 0, c)      # index 0 is never written
    lib.rc_pixel(w, h, p, -1, 1, c)     # x = -1 wraps to the end of row 0
    lib.rc_pixel(w, h, p, w, 1, c)      # x = W wraps to the start of row 2
    lib.rc_pixel(w, h, p, 0, h, c)      # past the end: nothing
    lib.rc_pixel(w, h, p, -2, 0, c)     # negative index: nothing
    assert _written(buf) == {(w - 1, 0), (0, 2)}


def test_rect_closed_forms():
    lib = refcpu.load()
    w, h = 8, 6
    c = (ctypes.c_uint8 * 4)(9, 9, 9, 255)
    buf = _overlay_buf(w, h)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    lib.rc_rect(w, h, p, 2, 1, 3, 3, c)
    assert _written(buf) == {(x, y) for x in (2, 3, 4) for y in (1, 2, 3)}
    buf[:] = 0
    lib.rc_rect(w, h, p, w, 0, 3, 3, c)   # x >= W: returns before drawing
    lib.rc_rect(w, h, p, 0, h, 3, 3, c)   # y >= H
    assert _written(buf) == set()
    lib.rc_rect(w, h, p, -1, 2, 3, 3, c)  # x = -1 wraps into the previous row
    # columns -1, 0, 1 of rows 2..4: column -1 is the last pixel of the row above
    assert _written(buf) == {(w - 1, 1), (0, 2), (1, 2), (w - 1, 2), (0, 3), (1, 3),
                             (w - 1, 3), (0, 4), (1, 4)}


def test_line_closed_forms():
    lib = refcpu.load()
    w, h = 16, 12
    c = (ctypes.c_uint8 * 4)(5, 5, 5, 255)
    buf = _overlay_buf(w, h)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    lib.rc_line(w, h, p, 1, 2, 6, 2, c)   # horizontal, both ends included
    assert _written(buf) == {(x, 2) for x in range(1, 7)}
    buf[:] = 0
    lib.rc_line(w, h, p, 3, 9, 3, 4, c)   # vertical, drawn from the first end
    assert _written(buf) == {(3, y) for y in range(4, 10)}
    buf[:] = 0
    lib.rc_line(w, h, p, 0, 0, 8, 2, c)   # step 2/8 = 0.25 exactly: y = floor(i / 4)
    # (0, 0) is flat index 0, which Pixel never writes
    assert _written(buf) == {(i, i // 4) for i in range(1, 9)}
    buf[:] = 0
    lib.rc_line(w, h, p, 4, 4, 4, 4, c)   # a point: side 0, one pixel (then NaN steps)
    assert _written(buf) == {(4, 4)}
    buf[:] = 0
    lib.rc_line(w, h, p, 10, 10, 2, 2, c)  # 45 degrees backwards
    assert _written(buf) == {(i, i) for i in range(2, 11)}
