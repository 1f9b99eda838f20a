"""CPU checks of the shortcuts the SASA arrangement kernel takes in place of
origami's exact FP64 arithmetic (sasa_core.h, built here for the host).

Each shortcut must reach the same decision as the arithmetic it replaces
(arrangement.rs:100-228, geometry.rs:67-100); these tests drive them at and
around the tolerance thresholds where a mistake would show.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def host():
    d = tempfile.mkdtemp(prefix="og_sasa_host_")
    so = os.path.join(d, "libsasa_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-o", so, os.path.join(HERE, "host", "sasa_core_host.cpp")], check=True)
    L = C.CDLL(so)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.host_intersect_counts.argtypes = [C.c_int, dp, ip, ip]
    L.host_intersect_points.argtypes = [C.c_int, dp, dp]
    L.host_midpoints.argtypes = [C.c_int, dp, dp]
    L.host_cover_arcs.argtypes = [C.c_int, dp, C.POINTER(C.c_float), dp]
    return L


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def _pairs_near(rng, n, targets, sigma_range=(-0.999, 0.999)):
    """Cap pairs whose exact c^2 = (s1^2 s2^2 - (sigma - c1 c2)^2) / (1 - sigma^2)
    equals one of the targets."""
    n1 = _unit(rng, n)
    sig = rng.uniform(*sigma_range, size=n)
    w = _unit(rng, n)
    w -= n1 * np.sum(w * n1, axis=1)[:, None]
    w /= np.linalg.norm(w, axis=1)[:, None]
    n2 = n1 * sig[:, None] + w * np.sqrt(1 - sig ** 2)[:, None]
    c1 = rng.uniform(-0.95, 0.95, size=n)
    t = rng.choice(targets, size=n)
    D = 1 - sig ** 2
    disc = np.maximum(D * ((1 - c1 ** 2) - t), 0)
    c2 = sig * c1 + rng.choice([-1.0, 1.0], size=n) * np.sqrt(disc)
    ok = np.abs(c2) < 1
    return np.concatenate([n1, c1[:, None], n2, c2[:, None]], axis=1)[ok]


def test_fast_count_matches_exact(host):
    rng = np.random.default_rng(0)
    th = 1e-9
    targets = np.concatenate([
        th * np.array([-1.6, -1.5, -1.4, -1.1, -1.0, -0.9, -0.6, -0.5, -0.4, 0.0,
                       0.4, 0.5, 0.6, 0.9, 1.0, 1.1, 1.4, 1.5, 1.6]),
        [-1e-3, 1e-3, -0.5, 0.3]])
    caps = np.concatenate([
        _pairs_near(rng, 400000, targets),
        _pairs_near(rng, 100000, targets, sigma_range=(0.99, 0.9999)),   # near-parallel
        np.concatenate([_unit(rng, 100000), rng.uniform(-1, 1, (100000, 1)),
                        _unit(rng, 100000), rng.uniform(-1, 1, (100000, 1))], axis=1),
    ])
    caps = np.ascontiguousarray(caps)
    n = len(caps)
    a = np.zeros(n, np.int32)
    b = np.zeros(n, np.int32)
    host.host_intersect_counts(n, _ptr(caps, C.c_double), _ptr(a, C.c_int), _ptr(b, C.c_int))
    assert np.array_equal(a, b), np.nonzero(a != b)[0][:10]
    # the thresholds were actually exercised on both sides
    assert set(np.unique(a)) == {0, 1, 2}


def test_points_of_well_conditioned_pairs_stay_on_circle(host):
    # the fast skip test relies on intersect_circles placing its points on
    # circle 1 to 1e-12 in height and 1e-8 in radius (sin a >= 0.05) when
    # 1 - sigma^2 >= 1e-2, tangent points included
    rng = np.random.default_rng(1)
    targets = np.array([-0.9e-9, -0.5e-9, 0.0, 0.5e-9, 0.9e-9, 1.1e-9, 1e-6, 1e-3, 0.1, 0.5])
    caps = _pairs_near(rng, 300000, targets, sigma_range=(-0.99499, 0.99499))
    caps = caps[np.sqrt(1 - caps[:, 3] ** 2) >= 0.05]
    caps = np.ascontiguousarray(caps)
    n = len(caps)
    pts = np.zeros((n, 6))
    host.host_intersect_points(n, _ptr(caps, C.c_double), _ptr(pts, C.c_double))
    n1, c1 = caps[:, 0:3], caps[:, 3]
    s1 = np.sqrt(1 - c1 ** 2)
    checked = 0
    for P in (pts[:, 0:3], pts[:, 3:6]):
        have = ~np.isnan(P[:, 0])
        h = np.sum(P[have] * n1[have], axis=1)
        rn = np.linalg.norm(P[have] - n1[have] * h[:, None], axis=1)
        assert np.abs(h - c1[have]).max() < 1e-12
        assert np.abs(rn - s1[have]).max() < 1e-8
        checked += have.sum()
    assert checked > 100000


def _approx_midpoint(a, n, ccw):
    # the kernel's approx_midpoint: FP64 rotation with FP32 sine and cosine
    h = np.sum(a * n, axis=1)
    r = a - n * h[:, None]
    t = np.cross(n, r)
    d = (0.5 * ccw).astype(np.float32)
    sd = np.sin(d).astype(np.float64)
    cd = np.cos(d).astype(np.float64)
    return n * h[:, None] + r * cd[:, None] + t * sd[:, None]


def test_approximate_midpoint_within_margin(host):
    # a cap containing the approximate midpoint with margin 1e-4 must contain
    # the exact one: the two differ by far less than the margin
    rng = np.random.default_rng(2)
    n = 200000
    nv = _unit(rng, n)
    ca = rng.uniform(-0.999, 0.999, n)
    u = _unit(rng, n)
    u -= nv * np.sum(u * nv, axis=1)[:, None]
    u /= np.linalg.norm(u, axis=1)[:, None]
    a = nv * ca[:, None] + u * np.sqrt(1 - ca ** 2)[:, None]
    ccw = rng.uniform(1e-9, 2 * np.pi, n)
    ccw[:1000] = 2 * np.pi
    rows = np.ascontiguousarray(np.concatenate([a, nv, ca[:, None], ccw[:, None]], axis=1))
    exact = np.zeros((n, 3))
    host.host_midpoints(n, _ptr(rows, C.c_double), _ptr(exact, C.c_double))
    err = np.linalg.norm(_approx_midpoint(a, nv, ccw) - exact, axis=1).max()
    assert err < 1e-6
    # and so does every cap with that margin: check the implication directly
    o = _unit(rng, n)
    co = np.sum(o * _approx_midpoint(a, nv, ccw), axis=1) - rng.uniform(0, 2e-4, n)
    sure = np.sum(o * _approx_midpoint(a, nv, ccw), axis=1) > co + 1e-4
    assert np.all(np.sum(o[sure] * exact[sure], axis=1) > co[sure] + 1e-12)


def _cover_arcs(host, pairs):
    pairs = np.ascontiguousarray(pairs)
    n = len(pairs)
    arcs = np.zeros((n, 2), np.float32)
    bas = np.zeros((n, 6))
    host.host_cover_arcs(n, _ptr(pairs, C.c_double), _ptr(arcs, C.c_float), _ptr(bas, C.c_double))
    return arcs, bas


def test_cover_arcs_inside_exact_arcs(host):
    # every angle the cover arc claims (SA_COV_EPS inside its ends) has
    # containment value >= SA_COV_DELTA = 1e-3 on the exact circle (the
    # depth that lets the kernel leave a covered circle's transitions out)
    rng = np.random.default_rng(4)
    n = 200000
    n1, n2 = _unit(rng, n), _unit(rng, n)
    c1 = rng.uniform(-0.99, 0.99, n)
    c2 = rng.uniform(-0.99, 0.99, n)
    pairs = np.concatenate([n1, c1[:, None], n2, c2[:, None]], axis=1)
    arcs, bas = _cover_arcs(host, pairs)
    full = arcs[:, 1] >= 100.0
    part = (arcs[:, 1] > 0) & ~full
    assert part.sum() > 50000 and full.sum() > 1000
    s1 = np.sqrt(1 - c1 ** 2)
    for sel, frac in ((part, np.linspace(0, 1, 9)), (full, np.linspace(0, 1, 9))):
        idx = np.nonzero(sel)[0]
        for f in frac:
            if sel is full:
                phi = 2 * np.pi * f * np.ones(len(idx))
            else:
                a = arcs[idx, 0].astype(np.float64)
                w2 = arcs[idx, 1].astype(np.float64)
                phi = a + 1e-5 + f * (w2 - 2e-5)
            X = (n1[idx] * c1[idx, None] + s1[idx, None]
                 * (np.cos(phi)[:, None] * bas[idx, 0:3] + np.sin(phi)[:, None] * bas[idx, 3:6]))
            val = np.sum(X * n2[idx], axis=1) - c2[idx]
            assert val.min() > 1e-3 - 1e-9, (f, val.min())


def _cover_walk(arcs):
    """The kernel's greedy cover walk (FP32) over one circle's arcs."""
    eps = np.float32(1e-5)
    two_pi = np.float32(2 * np.pi)
    if np.any(arcs[:, 1] >= 100.0):
        return True
    w = arcs[:, 1]
    if not np.any(w > 0):
        return False
    i1 = int(np.argmax(w))
    start = np.float32(arcs[i1, 0] + w[i1])
    length = np.float32(two_pi - w[i1])
    t = np.float32(0)
    for _ in range(len(arcs) + 2):
        p = np.float32(start + t)
        d = (p - arcs[:, 0]).astype(np.float32)
        d = d - two_pi * np.floor(d * (np.float32(1) / two_pi))
        ext = np.where((w > 0) & (d > eps) & (d < w - eps), w - d, 0).astype(np.float32)
        best = ext.max()
        if not best > 0:
            return False
        t = np.float32(t + best)
        if t >= length + eps:
            return True
    return False


def _cover_ext(d, ey, inverse):
    """The walk's extension from reduced offset d = p - start of an arc."""
    eps = np.float32(1e-5)
    two_pi = np.float32(2 * np.pi)
    q = d * (np.float32(1) / two_pi) if inverse else d / two_pi
    d = d - two_pi * np.floor(q)
    return np.where((ey > 0) & (d > eps) & (d < ey - eps), ey - d, 0).astype(np.float32)


def test_cover_turn_count_by_reciprocal():
    # The walk counts whole turns with floor(d * (1 / 2 pi)) instead of the
    # rounded quotient d / (2 pi).  The two quotients are within 2 ulp, so the
    # floors can differ only where d lies within a few 1e-6 of a multiple of
    # 2 pi; there the reduced angle is near 0 or near 2 pi and both forms give
    # no extension, since arc widths stay below 2 pi - 0.08.  Every float32 d
    # within 1e-4 of the multiples in the walk's range (d in [-3 pi, 7 pi])
    # against a grid of widths, and uniform samples elsewhere.
    two_pi = np.float32(2 * np.pi)
    inv = np.float32(1) / two_pi
    ds = []
    for k in (-1, 1, 2, 3):
        c = np.float32(k * 2 * np.pi)
        x, hi = np.float32(c - np.float32(1e-4)), np.float32(c + np.float32(1e-4))
        while x <= hi:
            ds.append(x)
            x = np.nextafter(x, np.float32(np.inf))
    d = np.array(ds, dtype=np.float32)[:, None]
    ey = np.concatenate([np.linspace(1e-3, 2 * np.pi - 0.09, 61),
                         [2 * np.pi - 0.0894]]).astype(np.float32)[None, :]
    assert np.array_equal(_cover_ext(d, ey, False), _cover_ext(d, ey, True))
    rng = np.random.default_rng(12)
    u = rng.uniform(-3 * np.pi, 7 * np.pi, 2_000_000).astype(np.float32)
    far = np.abs(u / (2 * np.pi) - np.round(u / (2 * np.pi))) * 2 * np.pi > 1e-4
    assert np.array_equal(np.floor(u[far] / two_pi), np.floor(u[far] * inv))
    e = rng.uniform(1e-3, 2 * np.pi - 0.09, u.size).astype(np.float32)
    assert np.array_equal(_cover_ext(u, e, False), _cover_ext(u, e, True))


def test_cover_walk_on_chain_atoms(host):
    # atoms of a jittered chain: every circle the walk declares covered has
    # each of 4001 sampled points inside another cap with the 1e-3 margin
    import math
    from origami_amd import system as S
    s = S.build_system(["Lys", "Trp", "Ala", "Leu", "Phe", "Glu"])
    rng = np.random.default_rng(6)
    pos = s.pos + rng.normal(scale=0.05, size=s.pos.shape)
    R = s.sasa_radius
    covered_n = checked = 0
    for i in np.nonzero(s.sasa_gamma != 0)[0]:
        d = np.linalg.norm(pos - pos[i], axis=1)
        caps = []
        for j in range(len(pos)):
            if j == i or d[j] > R[i] + R[j] or d[j] + R[j] <= R[i] or d[j] + R[i] <= R[j]:
                continue
            ca = (d[j] ** 2 + R[i] ** 2 - R[j] ** 2) / (2 * d[j] * R[i])
            caps.append(np.concatenate([(pos[j] - pos[i]) / d[j], [ca]]))
        caps = np.array(caps)
        nc = len(caps)
        for k in range(nc):
            if 1 - caps[k, 3] ** 2 < 0.0025:
                continue
            others = np.array([m for m in range(nc) if m != k])
            if len(others) == 0:
                continue
            pairs = np.concatenate([np.repeat(caps[k:k + 1], len(others), 0), caps[others]], axis=1)
            arcs, bas = _cover_arcs(host, pairs)
            if not _cover_walk(arcs):
                continue
            covered_n += 1
            u, v = bas[0, 0:3], bas[0, 3:6]
            sk = math.sqrt(1 - caps[k, 3] ** 2)
            phi = np.linspace(0, 2 * np.pi, 4001)
            X = caps[k, :3] * caps[k, 3] + sk * (np.cos(phi)[:, None] * u + np.sin(phi)[:, None] * v)
            val = (X @ caps[others, :3].T - caps[others, 3]).max(axis=1)
            assert val.min() > 1e-3 - 1e-9
            checked += 1
    assert checked > 200


def test_jvp_split_is_bit_identical(host):
    # the terms kernel forms each arc's / vertex's primal geometry once and
    # one tangent per direction; every term must equal the full JVP's bits
    host.host_jvp_split.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 5
    rng = np.random.default_rng(8)
    n = 20000
    pi = np.zeros((n, 3))
    dirs = [_unit(rng, n) for _ in range(3)]
    rad = rng.uniform(2.6, 3.2, (n, 4))
    dist = rng.uniform(1.0, 5.5, (n, 3))
    pts = [pi] + [pi + dirs[q] * dist[:, q:q + 1] for q in range(3)]
    rows = np.ascontiguousarray(np.concatenate(pts + [rad, rng.choice([-1.0, 1.0], (n, 2))], axis=1))
    af, asp = np.zeros((n, 12)), np.zeros((n, 12))
    vf, vsp = np.zeros((n, 9)), np.zeros((n, 9))
    host.host_jvp_split(n, _ptr(rows, C.c_double), _ptr(af, C.c_double), _ptr(asp, C.c_double),
                        _ptr(vf, C.c_double), _ptr(vsp, C.c_double))
    assert np.count_nonzero(af) > n and np.count_nonzero(vf) > n
    assert np.array_equal(af.view(np.int64), asp.view(np.int64))
    assert np.array_equal(vf.view(np.int64), vsp.view(np.int64))


def test_jvp_translation_invariance(host):
    # the terms kernel forms role i's tangent (arc: i, k, l, m; vertex: i, k,
    # l) as minus the sum of the other roles' tangents, since the arc angle
    # and the vertex angle depend on the positions only through differences;
    # pinned on the full JVPs, atom i away from the origin
    host.host_jvp_split.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 5
    rng = np.random.default_rng(9)
    n = 20000
    pi = rng.uniform(-40.0, 40.0, (n, 3))
    dirs = [_unit(rng, n) for _ in range(3)]
    rad = rng.uniform(2.6, 3.2, (n, 4))
    dist = rng.uniform(1.0, 5.5, (n, 3))
    pts = [pi] + [pi + dirs[q] * dist[:, q:q + 1] for q in range(3)]
    rows = np.ascontiguousarray(np.concatenate(pts + [rad, rng.choice([-1.0, 1.0], (n, 2))], axis=1))
    af, asp = np.zeros((n, 12)), np.zeros((n, 12))
    vf, vsp = np.zeros((n, 9)), np.zeros((n, 9))
    host.host_jvp_split(n, _ptr(rows, C.c_double), _ptr(af, C.c_double), _ptr(asp, C.c_double),
                        _ptr(vf, C.c_double), _ptr(vsp, C.c_double))
    for t, nr in ((af, 4), (vf, 3)):
        t = t.reshape(n, nr, 3)
        live = np.abs(t).max(axis=(1, 2)) > 0
        assert live.sum() > 2000
        scale = np.abs(t).max(axis=1)
        resid = np.abs(t[:, 0] + t[:, 1:].sum(axis=1))
        assert np.all(resid <= 1e-10 * scale + 1e-14), resid.max()
        # the negated sum, as the kernel forms it, against the direct tangent
        direct, formed = t[:, 0], -t[:, 1:].sum(axis=1)
        assert np.allclose(formed, direct, rtol=1e-9, atol=1e-13)


def test_pair_index_division_exact(host):
    # the arrangement kernel decodes (circle, other cap) pair indices with a
    # reciprocal multiply; it must equal integer division everywhere it is used
    host.host_pair_div_mismatches.restype = C.c_longlong
    assert host.host_pair_div_mismatches() == 0
