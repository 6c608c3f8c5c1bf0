"""Seeded initial conditions (host code of the library; no GPU)."""
import numpy as np


def test_component_counts_masses_and_types(gcs):
    st = gcs.collision(8192, 1001)
    assert len(st) == 16384
    for g, (lo, visible_type) in enumerate([(0, 0), (8192, 2)]):
        gal = st[lo:lo + 8192]
        halo, disc, bulge = gal[:1639], gal[1639:1639 + 5734], gal[1639 + 5734:]
        assert len(bulge) == 819
        assert np.all(halo["type"] == 1) and np.all(halo["r"] == 0)
        assert np.all(disc["type"] == visible_type) and np.all(bulge["type"] == visible_type)
        assert np.isclose(halo["mass"].astype(np.float64).sum(), 5.0, rtol=1e-5)
        assert np.isclose(disc["mass"].astype(np.float64).sum(), 1.0, rtol=1e-5)
        assert np.isclose(bulge["mass"].astype(np.float64).sum(), 0.25, rtol=1e-5)
        assert np.ptp(disc["mass"]) == 0 and np.ptp(bulge["mass"]) == 0
    assert tuple(st[1639][k] for k in "rgb") == (255, 100, 50)
    assert tuple(st[8192 + 1639][k] for k in "rgb") == (50, 100, 255)


def test_deterministic_and_seed_dependent(gcs):
    a = gcs.collision(3000, 5)
    b = gcs.collision(3000, 5)
    c = gcs.collision(3000, 6)
    assert a.tobytes() == b.tobytes()
    assert a.tobytes() != c.tobytes()


def test_geometry(gcs):
    st = gcs.collision(20000, 3)
    a, b = st[:20000], st[20000:]
    ra = np.sqrt(a["x"].astype(np.float64) ** 2 + a["y"] ** 2 + a["z"] ** 2)
    assert ra.max() < 60.0 + 1e-3
    disc = a[4000:4000 + 14000]
    R = np.sqrt(disc["x"].astype(np.float64) ** 2 + disc["y"] ** 2)
    assert R.max() <= 30.0 and np.abs(disc["z"]).max() <= 3.0
    # The disc is thin: Miyamoto-Nagai b/a = 0.1.
    assert np.median(np.abs(disc["z"])) < 0.5 < np.median(R)
    # Galaxy B centre of mass near (40, 10, 0), moving (-0.5, 0, 0).
    m = b["mass"].astype(np.float64)
    com = np.array([(b[k] * m).sum() / m.sum() for k in ("x", "y", "z")])
    vcm = np.array([(b[k] * m).sum() / m.sum() for k in ("vx", "vy", "vz")])
    assert np.allclose(com, [40, 10, 0], atol=1.5)
    assert np.allclose(vcm, [-0.5, 0, 0], atol=0.1)
    # Disc of B tilted 30 degrees about x: angular momentum direction.
    d = b[4000:4000 + 14000]
    x, y, z = (d[k].astype(np.float64) - c for k, c in zip("xyz", com))
    vx, vy, vz = (d[k].astype(np.float64) - c for k, c in zip(("vx", "vy", "vz"), vcm))
    L = np.array([(y * vz - z * vy).sum(), (z * vx - x * vz).sum(), (x * vy - y * vx).sum()])
    L /= np.linalg.norm(L)
    assert np.allclose(L, [0, -np.sin(np.radians(30)), np.cos(np.radians(30))], atol=0.05)


def test_circular_speed_from_enclosed_mass(gcs):
    st = gcs.collision(20000, 4)
    a = st[:20000]
    r = np.sqrt(a["x"].astype(np.float64) ** 2 + a["y"] ** 2 + a["z"] ** 2)
    v = np.sqrt(a["vx"].astype(np.float64) ** 2 + a["vy"] ** 2 + a["vz"] ** 2)
    order = np.argsort(r, kind="stable")
    menc = np.empty_like(r)
    menc[order] = np.concatenate([[0], np.cumsum(a["mass"][order].astype(np.float64))[:-1]])
    vc = np.sqrt(menc / r)
    ratio = v[r > 1] / vc[r > 1]
    assert abs(np.median(ratio) - 1.0) < 0.01  # 5% Gaussian per component


def test_random_specs(gcs):
    for k in range(50):
        s = gcs.random_spec(77, k, 4096, 65536)
        n = s.n_a + s.n_b
        assert 4096 <= n <= 65536 and abs(s.n_a - s.n_b) <= 1
        d = np.linalg.norm(list(s.offset))
        assert 20 - 1e-4 <= d <= 60 + 1e-4
        sp = np.linalg.norm(list(s.velocity))
        assert 0.2 - 1e-5 <= sp <= 0.8 + 1e-5
        assert np.dot(list(s.offset), list(s.velocity)) < 0
        assert 0 <= s.tilt_deg <= 90
