"""CPU: pin the arm pipeline's checker (oracle/arm_ref.py) by closed forms.

Forward kinematics: all joints at zero lays the links along +x at the
cumulative lengths; a lone rotation of joint 0 (about z) or joint 1 (about
y) turns the rest of the chain rigidly about that joint; consecutive joints
stay exactly one link length apart for random configurations.  Pipeline: a
box placed across a link is a collision, a chain far from every box has no
candidate pair.
"""
import numpy as np

import arm_ref as A
import pyoracle
from mfcl import workloads as W


def test_fk_zero_configuration():
    lens = np.array([0.3, 0.2, 0.25], np.float32)
    pose = A.fk(np.zeros((1, 3)), lens)[0].astype(np.float64)
    ends = np.cumsum(lens.astype(np.float64))
    np.testing.assert_allclose(pose[:, 0], ends - lens / 2, atol=1e-7)
    np.testing.assert_allclose(pose[:, 1:3], 0, atol=0)
    np.testing.assert_allclose(pose[:, 3:7], [[1, 0, 0, 0]] * 3, atol=0)


def test_fk_single_joint_rotations():
    lens = np.array([0.3, 0.2, 0.25, 0.15], np.float32).astype(np.float64)
    th = 0.7
    # joint 0 about z: the whole chain turns about the origin
    p = A.fk(np.array([[th, 0, 0, 0]]), lens)[0].astype(np.float64)
    s = np.cumsum(lens) - lens / 2
    np.testing.assert_allclose(p[:, 0], np.cos(th) * s, atol=1e-6)
    np.testing.assert_allclose(p[:, 1], np.sin(th) * s, atol=1e-6)
    np.testing.assert_allclose(p[:, 3:7], [[np.cos(th / 2), 0, 0, np.sin(th / 2)]] * 4, atol=1e-7)
    # joint 1 about y: links 1.. turn about the joint at (L0, 0, 0); x maps to (cos, 0, -sin)
    p = A.fk(np.array([[0, th, 0, 0]]), lens)[0].astype(np.float64)
    d = np.cumsum(lens[1:]) - lens[1:] / 2
    np.testing.assert_allclose(p[0, :3], [lens[0] / 2, 0, 0], atol=1e-7)
    np.testing.assert_allclose(p[1:, 0], lens[0] + np.cos(th) * d, atol=1e-6)
    np.testing.assert_allclose(p[1:, 2], -np.sin(th) * d, atol=1e-6)


def test_fk_rigid_links():
    rng = np.random.default_rng(0)
    lens = rng.uniform(0.1, 0.4, size=7).astype(np.float32)
    q = rng.uniform(-np.pi, np.pi, size=(200, 7))
    p = A.fk(q, lens).astype(np.float64)
    w, x, y, z = (p[..., k] for k in range(3, 7))
    ax = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)], -1)
    start = p[..., :3] - ax * (lens[None, :, None] / 2)
    end = p[..., :3] + ax * (lens[None, :, None] / 2)
    np.testing.assert_allclose(start[:, 0], 0, atol=1e-6)             # the base is the origin
    np.testing.assert_allclose(start[:, 1:], end[:, :-1], atol=2e-6)  # links meet at the joints
    np.testing.assert_allclose(np.linalg.norm(p[..., 3:7], axis=-1), 1, atol=1e-6)


def test_pipeline_hit_and_clear(oracle):
    scene = W.arm_scene(4, seed=6, n_links=3, n_boxes=2)
    # box 0 across link 0 of a straight arm, box 1 far away
    scene["box_pose"][0, :3] = [scene["link_len"][0] / 2, 0, 0]
    scene["box_pose"][1, :3] = [50, 50, 50]
    q = np.zeros((2, 3), np.float32)
    q[1, 0] = np.pi  # arm folded back along -x: clear of box 0 by more than the margin
    scene["box_pose"][0, 3:] = [1, 0, 0, 0]
    scene["params"][3] = [0.05, 0.05, 0.05, 0]
    scene["box_half"][0] = 0.025
    S = A.shape_set(scene)
    mn, mx = W.box_aabbs(scene["box_pose"], scene["box_half"])
    tree = pyoracle.AABBTree(mn, mx, np.arange(2))
    flags, mind, npairs = A.pipeline(scene, q, 0.05, 8, S, tree)
    assert flags[0] & 1 and mind[0] < 0
    assert flags[1] == 0 and np.isinf(mind[1])


# ---- heightmap ---------------------------------------------------------------
def _hm_scene(n=256):
    return W.arm_scene(n, seed=6, heightmap=True)


def test_fbm_heightmap_shape_and_range():
    h = W.fbm_heightmap(np.random.default_rng(0))
    assert h.shape == (256, 256) and h.dtype == np.uint16
    assert h.min() == 0 and h.max() == 100  # normalised to [0, 0.1] m
    # smooth at the pixel scale: neighbouring pixels differ by a few mm
    assert np.abs(np.diff(h.astype(np.int64), axis=1)).max() <= 10


def test_hm_candidates_cover_every_colliding_pixel():
    """The flat algorithm only culls: for links with candidates, every pixel
    box of the whole map that the oracle finds colliding (brute force over
    all 65,536 pixels) is a candidate."""
    sc = _hm_scene(64)
    pose = A.fk(sc["q"], sc["link_len"])
    k, x, y, *_ = A.hm_candidates(sc, pose)
    S = A.shape_set(sc)
    fy, fx = sc["hm_heights"].shape
    qs = np.unique(k)[:6]
    assert len(qs) >= 3
    for q in qs:
        allk = np.full(fx * fy, q)
        ax, ay = np.meshgrid(np.arange(fx), np.arange(fy))
        ax, ay = ax.reshape(-1), ay.reshape(-1)
        keep = sc["hm_heights"][ay, ax] > 0
        pa, pb, qa, qb = A.hm_pairs(sc, allk[keep], ax[keep], ay[keep], pose)
        ref, _ = S.collide(pa, pb, pyoracle.pose12(qa), pyoracle.pose12(qb), max_contacts=1, penetration=False,
                           threads=8)
        hit_pix = set(zip(ax[keep][ref["hit"] == 1].tolist(), ay[keep][ref["hit"] == 1].tolist()))
        cand = set(zip(x[k == q].tolist(), y[k == q].tolist()))
        assert hit_pix <= cand, (q, len(hit_pix - cand))


def test_hm_pixel_box_and_roi_closed_forms():
    """Pixel boxes: sides (res, res, h) and world centres on the map's
    rotated grid; a link placed inside a tall pixel's box collides, one
    raised above the map's highest pixel has no candidate."""
    sc = _hm_scene(4)
    H = sc["hm_heights"]
    res = sc["hm_res"]
    y, x = np.unravel_index(np.argmax(H), H.shape)
    pa, pb, qa, qb = A.hm_pairs(sc, np.array([0]), np.array([x]), np.array([y]), A.fk(sc["q"], sc["link_len"]))
    ang = 0.3
    cx, cy = (x - 128 + 0.5) * res, (y - 128 + 0.5) * res
    want = np.array([0.3 + np.cos(ang) * cx - np.sin(ang) * cy, np.sin(ang) * cx + np.cos(ang) * cy,
                     -0.45 + 0.5 * H[y, x] * 1e-3])
    np.testing.assert_allclose(qa[0, :3], want, atol=2e-7)
    assert pa[0] == sc["hm_box_shape0"] + y * 256 + x
    np.testing.assert_allclose(sc["params"][pa[0], :3], [res, res, H[y, x] * 1e-3], rtol=1e-7)
    # link 0 of configuration 0 moved: centred inside that pixel box / far above the map
    for z, expect in ((0.5 * H[y, x] * 1e-3, True), (0.5, False)):
        p = A.fk(sc["q"][:1], sc["link_len"]).copy()
        ctr = np.array([cx, cy, z])
        p[0, 0, :3] = np.array([0.3 + np.cos(ang) * ctr[0] - np.sin(ang) * ctr[1],
                                np.sin(ang) * ctr[0] + np.cos(ang) * ctr[1], -0.45 + ctr[2]], np.float32)
        hit, cnt, _ = A.hm_collide(sc, A.shape_set(sc), p, threads=1)
        assert hit[0] == expect and (cnt[0] > 0) == expect
