"""Seeded synthetic multi-view scene generator.

Stands in for the MVSNet-style input folders the reference reads
(`images/%08d.jpg`, `cams/%08d_cam.txt`, `pair.txt`; main.cpp:336-341), since
no dataset is available.  A scene is 8 planar rectangles (one background plane
plus 7 patches) with depths in [2, 6] in the reference camera and normals
within 60 degrees of its viewing axis.  Planes carry a 4-octave value-noise
texture (intensity 0..255); 20 % of the reference view shows weak-texture
regions: constant intensity plus independent N(0, 2^2) noise in every view.
Cameras sit on an arc around the scene centre, all looking at it.  Ground-truth
depth (camera z) and world normals are kept for every view.
"""
from dataclasses import dataclass, field
import numpy as np


@dataclass
class View:
    K: np.ndarray          # 3x3 float32
    R: np.ndarray          # 3x3 float32, world -> camera
    t: np.ndarray          # 3 float32
    image: np.ndarray      # HxW uint8
    depth: np.ndarray      # HxW float32 ground-truth camera z
    normal: np.ndarray     # HxWx3 float32 world-frame normal (facing the camera)
    weak: np.ndarray       # HxW bool, pixel sees a weak-texture region


@dataclass
class Scene:
    width: int
    height: int
    views: list
    pairs: list            # pairs[i] = list of source view indices for reference i
    depth_min: float = 2.0
    depth_max: float = 6.0
    meta: dict = field(default_factory=dict)


def _value_noise_tables(rng, n_octaves, base_cells):
    tabs = []
    for o in range(n_octaves):
        cells = base_cells * (2 ** o)
        tabs.append(rng.random((cells + 2, cells + 2)).astype(np.float32))
    return tabs


def _value_noise(tabs, u, v, extent):
    """Sum of octaves of smooth value noise at plane coordinates (u, v) in [-extent, extent]."""
    out = np.zeros(u.shape, np.float32)
    amp_sum = 0.0
    amp = 1.0
    for tab in tabs:
        cells = tab.shape[0] - 2
        gu = (u / (2 * extent) + 0.5) * cells
        gv = (v / (2 * extent) + 0.5) * cells
        gu = np.clip(gu, 0, cells - 1e-3)
        gv = np.clip(gv, 0, cells - 1e-3)
        iu = np.floor(gu).astype(np.int64)
        iv = np.floor(gv).astype(np.int64)
        fu = gu - iu
        fv = gv - iv
        su = fu * fu * (3 - 2 * fu)
        sv = fv * fv * (3 - 2 * fv)
        a = tab[iv, iu] + su * (tab[iv, iu + 1] - tab[iv, iu])
        b = tab[iv + 1, iu] + su * (tab[iv + 1, iu + 1] - tab[iv + 1, iu])
        out += amp * (a + sv * (b - a))
        amp_sum += amp
        amp *= 0.75
    return out / amp_sum


def _look_at(center, target):
    z = target - center
    z = z / np.linalg.norm(z)
    up = np.array([0.0, 1.0, 0.0])
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])          # rows: camera axes in world
    t = -R @ center
    return R, t


def make_scene(width=320, height=240, focal=300.0, n_views=5, arc_deg=30.0, seed=0x5EED0001,
               n_planes=8, weak_fraction=0.2, n_src=None, texture_px=20.0, supersample=2):
    """Render a seeded scene.  `texture_px` is the wavelength of the coarsest
    texture octave in pixels at depth 4, so the texture statistics in pixel
    units do not depend on the resolution."""
    rng = np.random.default_rng(seed)
    W, H = int(width), int(height)
    f = float(focal)
    K = np.array([[f, 0, (W - 1) / 2.0], [0, f, (H - 1) / 2.0], [0, 0, 1]], np.float64)
    target = np.array([0.0, 0.0, 4.0])
    radius = 4.0
    angles = np.deg2rad(np.linspace(-arc_deg / 2, arc_deg / 2, n_views))
    ref_mid = n_views // 2
    cams = []
    for a in angles:
        c = target + radius * np.array([np.sin(a), 0.0, -np.cos(a)])
        R, t = _look_at(c, target)
        cams.append((R, t, c))

    # planes, defined in the frame of the middle camera (== world at angle 0)
    half_w = 4.0 * (W / 2) / f * 1.6
    half_h = 4.0 * (H / 2) / f * 1.6
    planes = []
    # background: depth ~5.5, mild tilt
    def rand_normal(max_deg):
        while True:
            th = np.deg2rad(rng.uniform(0, max_deg))
            ph = rng.uniform(0, 2 * np.pi)
            n = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)])
            return n
    bg_n = rand_normal(20.0)
    bg_c = np.array([0.0, 0.0, 5.6])
    planes.append(dict(c=bg_c, n=bg_n, hw=half_w * 2.2, hh=half_h * 2.2))
    for k in range(n_planes - 1):
        depth = rng.uniform(2.6, 5.0)
        px = rng.uniform(-0.8, 0.8) * (W / 2) / f * depth
        py = rng.uniform(-0.8, 0.8) * (H / 2) / f * depth
        n = rand_normal(55.0)
        size = rng.uniform(0.18, 0.42) * depth * (W / f)
        planes.append(dict(c=np.array([px, py, depth]), n=n, hw=size * rng.uniform(0.6, 1.0),
                           hh=size * rng.uniform(0.6, 1.0)))
    extent = max(half_w, half_h) * 2.4
    # texture tables per plane: base cell count so the coarsest octave has
    # wavelength texture_px pixels at depth 4
    base_wave = texture_px * 4.0 / f
    base_cells = max(2, int(np.ceil(2 * extent / base_wave)))
    for pl in planes:
        n = pl["n"] / np.linalg.norm(pl["n"])
        a1 = np.cross(n, [0.0, 1.0, 0.0])
        a1 /= np.linalg.norm(a1)
        a2 = np.cross(n, a1)
        pl.update(n=n, a1=a1, a2=a2)
        pl["tex"] = _value_noise_tables(rng, 4, base_cells)
        pl["weak_tab"] = _value_noise_tables(rng, 1, max(2, int(2 * extent / 1.2)))
        pl["weak_level"] = rng.uniform(60, 200)
        pl["gain"] = rng.uniform(0.8, 1.2)

    ss = int(supersample)

    def render(R, t, c, noise_rng):
        # rays through (x, y) with supersampling offsets
        xs = (np.arange(W * ss) + 0.5) / ss - 0.5
        ys = (np.arange(H * ss) + 0.5) / ss - 0.5
        X, Y = np.meshgrid(xs, ys)
        Kinv = np.linalg.inv(K)
        dcam = np.stack([Kinv[0, 0] * X + Kinv[0, 2], Kinv[1, 1] * Y + Kinv[1, 2], np.ones_like(X)], -1)
        dworld = dcam @ R          # R^T d
        best_t = np.full(X.shape, np.inf)
        best_k = np.full(X.shape, -1, np.int64)
        best_u = np.zeros(X.shape)
        best_v = np.zeros(X.shape)
        for k, pl in enumerate(planes):
            denom = dworld @ pl["n"]
            tt = ((pl["c"] - c) @ pl["n"]) / np.where(np.abs(denom) < 1e-12, 1e-12, denom)
            hit = c + tt[..., None] * dworld
            rel = hit - pl["c"]
            u = rel @ pl["a1"]
            v = rel @ pl["a2"]
            ok = (tt > 0) & (np.abs(u) <= pl["hw"]) & (np.abs(v) <= pl["hh"]) & (tt < best_t)
            best_t = np.where(ok, tt, best_t)
            best_k = np.where(ok, k, best_k)
            best_u = np.where(ok, u, best_u)
            best_v = np.where(ok, v, best_v)
        inten = np.zeros(X.shape, np.float32)
        weakm = np.zeros(X.shape, bool)
        for k, pl in enumerate(planes):
            m = best_k == k
            if not m.any():
                continue
            u = best_u[m]
            v = best_v[m]
            tex = _value_noise(pl["tex"], u, v, extent)
            val = np.clip((tex - 0.5) * 2.2 * pl["gain"] + 0.5, 0, 1) * 255.0
            wk = _value_noise(pl["weak_tab"], u, v, extent) > pl["weak_thr"]
            val = np.where(wk, pl["weak_level"], val)
            inten[m] = val
            weakm[m] = wk
        # depth along the camera z axis
        depth_s = best_t * (dworld @ R[2])
        img = inten.reshape(H, ss, W, ss).mean(axis=(1, 3))
        weak_px = weakm.reshape(H, ss, W, ss).mean(axis=(1, 3)) > 0.5
        img = img + np.where(weak_px, noise_rng.normal(0.0, 2.0, img.shape), 0.0)
        img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
        # ground truth at the pixel centre sample
        ci = ss // 2
        depth = depth_s[ci::ss, ci::ss].astype(np.float32)
        kk = best_k[ci::ss, ci::ss]
        normal = np.zeros((H, W, 3), np.float32)
        for k, pl in enumerate(planes):
            m = kk == k
            n = pl["n"]
            # orient towards the camera
            nn = n if np.dot(n, c - pl["c"]) > 0 else -n
            normal[m] = nn
        return img, depth, normal, weak_px

    # choose each plane's weak threshold so ~weak_fraction of the middle view is weak
    R0, t0, c0 = cams[ref_mid]
    for pl in planes:
        pl["weak_thr"] = 2.0  # no weak first
    # sample weak-noise values over the middle view to set one common quantile
    xs = np.arange(W) - 0.0
    ys = np.arange(H) - 0.0
    X, Y = np.meshgrid(xs, ys)
    Kinv = np.linalg.inv(K)
    dcam = np.stack([Kinv[0, 0] * X + Kinv[0, 2], Kinv[1, 1] * Y + Kinv[1, 2], np.ones_like(X)], -1)
    dworld = dcam @ R0
    best_t = np.full(X.shape, np.inf)
    wv = np.zeros(X.shape, np.float32)
    for k, pl in enumerate(planes):
        denom = dworld @ pl["n"]
        tt = ((pl["c"] - c0) @ pl["n"]) / np.where(np.abs(denom) < 1e-12, 1e-12, denom)
        hit = c0 + tt[..., None] * dworld
        rel = hit - pl["c"]
        u = rel @ pl["a1"]
        v = rel @ pl["a2"]
        ok = (tt > 0) & (np.abs(u) <= pl["hw"]) & (np.abs(v) <= pl["hh"]) & (tt < best_t)
        best_t = np.where(ok, tt, best_t)
        wv = np.where(ok, _value_noise(pl["weak_tab"], u, v, extent), wv)
    thr = float(np.quantile(wv, 1.0 - weak_fraction))
    for pl in planes:
        pl["weak_thr"] = thr

    noise_rng = np.random.default_rng(seed + 1)
    views = []
    for (R, t, c) in cams:
        img, depth, normal, weak = render(R, t, c, noise_rng)
        views.append(View(K=K.astype(np.float32), R=R.astype(np.float32), t=t.astype(np.float32), image=img,
                          depth=depth, normal=normal, weak=weak))
    centers = np.array([c for (_, _, c) in cams])
    pairs = []
    ns = (n_views - 1) if n_src is None else int(n_src)
    for i in range(n_views):
        d = np.linalg.norm(centers - centers[i], axis=1)
        order = [j for j in np.argsort(d, kind="stable") if j != i][:ns]
        pairs.append([int(j) for j in order])
    return Scene(width=W, height=H, views=views, pairs=pairs, depth_min=2.0, depth_max=6.0,
                 meta=dict(seed=seed, focal=f, arc_deg=arc_deg, n_views=n_views, ref_mid=ref_mid))


def config1(seed=0x5EED0001):
    """1 reference + 4 sources, 320x240, f=300, 5 cameras on a 30 degree arc."""
    return make_scene(320, 240, 300.0, 5, 30.0, seed=seed)


def config2(seed=0x5EED0001, n_views=20):
    """1600x1200, f=1500, 20 cameras on a 60 degree arc, 10 nearest sources each."""
    return make_scene(1600, 1200, 1500.0, n_views, 60.0, seed=seed, n_src=10, supersample=1)


def write_folder(scene, folder, views=None):
    """Write the scene in the reference's input layout (main.cpp:336-341):
    images/%08d.pgm (lossless 8-bit gray instead of JPEG), cams/%08d_cam.txt
    (extrinsic 4x4, intrinsic 3x3, depth_min interval depth_num depth_max, as
    ReadCamera parses them, DPE.cpp:341-382) and pair.txt (GenerateSampleList's
    format, main.cpp:102-145; score = 1 / (1 + distance between centres))."""
    import os
    views = list(range(len(scene.views))) if views is None else list(views)
    os.makedirs(os.path.join(folder, "images"), exist_ok=True)
    os.makedirs(os.path.join(folder, "cams"), exist_ok=True)
    f9 = lambda v: "%.9g" % float(np.float32(v))
    for i in views:
        v = scene.views[i]
        H, W = v.image.shape
        with open(os.path.join(folder, "images", "%08d.pgm" % i), "wb") as f:
            f.write(b"P5\n%d %d\n255\n" % (W, H))
            f.write(np.ascontiguousarray(v.image, np.uint8).tobytes())
        lines = ["extrinsic"]
        for r in range(3):
            lines.append(" ".join(f9(v.R[r, c]) for c in range(3)) + " " + f9(v.t[r]))
        lines += ["0 0 0 1", "", "intrinsic"]
        for r in range(3):
            lines.append(" ".join(f9(v.K[r, c]) for c in range(3)))
        lines += ["", "%s %s %d %s" % (f9(scene.depth_min), f9((scene.depth_max - scene.depth_min) / 256), 256,
                                        f9(scene.depth_max))]
        with open(os.path.join(folder, "cams", "%08d_cam.txt" % i), "w") as f:
            f.write("\n".join(lines) + "\n")
    centers = {i: -scene.views[i].R.astype(np.float64).T @ scene.views[i].t.astype(np.float64) for i in views}
    with open(os.path.join(folder, "pair.txt"), "w") as f:
        f.write("%d\n" % len(views))
        for i in views:
            srcs = [j for j in scene.pairs[i] if j in centers]
            f.write("%d\n" % i)
            f.write("%d %s\n" % (len(srcs), " ".join(
                "%d %.6f" % (j, 1.0 / (1.0 + np.linalg.norm(centers[i] - centers[j]))) for j in srcs)))
