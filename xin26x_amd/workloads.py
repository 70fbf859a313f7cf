"""Seeded synthetic workloads (10-bit 4:2:0) shared by tests, smoke() and bench.py.

Inputs are generated on the CPU with a counter-based integer RNG (splitmix64)
and Box-Muller Gaussians in float64, so every consumer sees identical samples.

  frame_pair_416x240()  - integer full search, 16x16 blocks, +/-16, lambda 4
  frame_pair_1080p()    - fractional refinement: per-CTU motion field, QTMT
                          partitions, 32 fractional candidates per block
  residual_blocks()     - forward transforms: uneven sizes and MTS types
"""
import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)

# VVC luma / chroma interpolation filters (ITU-T H.266 Tables 27-29), used
# only to synthesise motion-shifted current frames.
LUMA_FILTER = np.array([
    [0, 0, 0, 64, 0, 0, 0, 0], [0, 1, -3, 63, 4, -2, 1, 0], [-1, 2, -5, 62, 8, -3, 1, 0],
    [-1, 3, -8, 60, 13, -4, 1, 0], [-1, 4, -10, 58, 17, -5, 1, 0], [-1, 4, -11, 52, 26, -8, 3, -1],
    [-1, 3, -9, 47, 31, -10, 4, -1], [-1, 4, -11, 45, 34, -10, 4, -1], [-1, 4, -11, 40, 40, -11, 4, -1],
    [-1, 4, -10, 34, 45, -11, 4, -1], [-1, 4, -10, 31, 47, -9, 3, -1], [-1, 3, -8, 26, 52, -11, 4, -1],
    [0, 1, -5, 17, 58, -10, 4, -1], [0, 1, -4, 13, 60, -8, 3, -1], [0, 1, -3, 8, 62, -5, 2, -1],
    [0, 1, -2, 4, 63, -3, 1, 0]], dtype=np.int64)
CHROMA_FILTER = np.array([
    [0, 64, 0, 0], [-1, 63, 2, 0], [-2, 62, 4, 0], [-2, 60, 7, -1], [-2, 58, 10, -2],
    [-3, 57, 12, -2], [-4, 56, 14, -2], [-4, 55, 15, -2], [-4, 54, 16, -2], [-5, 53, 18, -2],
    [-6, 52, 20, -2], [-6, 49, 24, -3], [-6, 46, 28, -4], [-5, 44, 29, -4], [-4, 42, 30, -4],
    [-4, 39, 33, -4], [-4, 36, 36, -4], [-4, 33, 39, -4], [-4, 30, 42, -4], [-4, 29, 44, -5],
    [-4, 28, 46, -6], [-3, 24, 49, -6], [-2, 20, 52, -6], [-2, 18, 53, -5], [-2, 16, 54, -4],
    [-2, 15, 55, -4], [-2, 14, 56, -4], [-2, 12, 57, -3], [-2, 10, 58, -2], [-1, 7, 60, -2],
    [0, 4, 62, -2], [0, 2, 63, -1]], dtype=np.int64)

# The 32 fractional candidates of the batched refinement, 1/16 pel, around an
# integer-pel base MV: rings at 8 (half), 4 (quarter), 12 and 2 (1/16 steps),
# each in the neighbour order L, R, T, B, TL, TR, BL, BR of
# h266_encoder/source/h266_sub_pel_refine.c:47-55.
_RING = [(-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1)]
CAND32 = np.array([(dx * r, dy * r) for r in (8, 4, 12, 2) for (dx, dy) in _RING], dtype=np.int16)
CTU = 64
LAMBDA_REFINE = 8192            # sadLambda at QP 30, 10-bit: index 30 + 12 (h266_enc_init.c:86-105, :384)


def splitmix64(seed, n, stream=0):
    """n outputs of splitmix64 counter mode: z_i = mix(seed + (i+1)*golden)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = base + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, stream=0):
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def randint(seed, n, lo, hi, stream=0):
    """Integers in [lo, hi] inclusive."""
    return lo + (splitmix64(seed, n, stream) % np.uint64(hi - lo + 1)).astype(np.int64)


def gaussian(seed, n, sigma, stream=0):
    """Box-Muller in float64."""
    m = (n + 1) // 2
    u1 = uniform01(seed, m, stream * 2 + 1)
    u2 = uniform01(seed, m, stream * 2 + 2)
    u1 = np.maximum(u1, 1e-300)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
    return z * sigma


def _box7(a):
    p = np.pad(a, 3, mode="edge")
    c = np.cumsum(np.cumsum(p, axis=0), axis=1)
    c = np.pad(c, ((1, 0), (1, 0)))
    h, w = a.shape
    s = c[7:7 + h, 7:7 + w] - c[0:h, 7:7 + w] - c[7:7 + h, 0:w] + c[0:h, 0:w]
    return s / 49.0


def smooth_field(h, w, seed, stream=0):
    """Uniform noise low-passed by a 7x7 box filter 3 times, rescaled to [64, 960]."""
    a = uniform01(seed, h * w, stream).reshape(h, w)
    for _ in range(3):
        a = _box7(a)
    lo, hi = a.min(), a.max()
    return np.rint(64 + (a - lo) * (896.0 / max(hi - lo, 1e-12))).astype(np.int64)


def _interp(plane, x0, y0, w, h, mvx, mvy, taps, fbits):
    """VVC separable interpolation (PIXEL output) of a w x h block at (x0, y0)
    displaced by (mvx, mvy) in 1/2^fbits sample units; reads outside the plane
    take the nearest edge sample."""
    flt = LUMA_FILTER if taps == 8 else CHROMA_FILTER
    m = (1 << fbits) - 1
    fx, fy = mvx & m, mvy & m
    ix, iy = x0 + (mvx >> fbits), y0 + (mvy >> fbits)
    H, W = plane.shape
    ys = np.clip(np.arange(iy - taps // 2 + 1, iy + h + taps // 2), 0, H - 1)
    xs = np.clip(np.arange(ix - taps // 2 + 1, ix + w + taps // 2), 0, W - 1)
    reg = plane[np.ix_(ys, xs)].astype(np.int64)
    t = np.zeros((h + taps - 1, w), np.int64)
    for j in range(taps):
        t += flt[fx][j] * reg[:, j:j + w]
    t = (t - (8192 << 2)) >> 2
    s = np.zeros((h, w), np.int64)
    for k in range(taps):
        s += flt[fy][k] * t[k:k + h, :]
    return np.clip((s + 512 + (8192 << 6)) >> 10, 0, 1023)


def pad_plane(a, pad):
    return np.pad(a, pad, mode="edge")


# ---------------------------------------------------------------------------
# 416x240: integer full search
# ---------------------------------------------------------------------------
def frame_pair_416x240(seed=1):
    """Reference: smooth field; current: reference shifted by (+13, -7)
    quarter-pel through the 8-tap filter plus N(0, 4^2), clipped."""
    H, W = 240, 416
    ref = smooth_field(H, W, seed)
    cur = _interp(ref, 0, 0, W, H, 13 * 4, -7 * 4, 8, 4)
    cur = np.clip(np.rint(cur + gaussian(seed, H * W, 4.0, stream=7).reshape(H, W)), 0, 1023)
    return cur.astype(np.uint16), ref.astype(np.uint16)


# ---------------------------------------------------------------------------
# 1080p: fractional refinement
# ---------------------------------------------------------------------------
def qtmt_partition(ctu_x, ctu_y, fw, fh, seed, max_depth=4):
    """Seeded random QT/BT/TT split of one 64x64 CTU: block sides 4..64,
    aspect ratio at most 4, QT only before any multi-type split; blocks that
    cross the frame edge are split by QT until inside, blocks fully outside
    are dropped.  Returns a list of (x, y, w, h)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []

    def rec(x, y, w, h, depth, qt_ok):
        if x >= fw or y >= fh:
            return
        if x + w > fw or y + h > fh:
            hw, hh = w // 2, h // 2
            for (dx, dy) in ((0, 0), (hw, 0), (0, hh), (hw, hh)):
                rec(x + dx, y + dy, hw, hh, depth + 1, qt_ok)
            return
        opts = [("none",)]
        if depth < max_depth:
            if qt_ok and w == h and w >= 8:
                opts.append(("qt",))
            if h >= 8 and w <= 4 * (h // 2):
                opts.append(("bth",))
            if w >= 8 and h <= 4 * (w // 2):
                opts.append(("btv",))
            if h >= 16 and w <= 4 * (h // 4):
                opts.append(("tth",))
            if w >= 16 and h <= 4 * (w // 4):
                opts.append(("ttv",))
        # favour splitting large blocks, stopping small ones
        weights = np.array([0.15 + 0.1 * depth] + [1.0] * (len(opts) - 1))
        if w * h >= 1024 and len(opts) > 1:
            weights[0] = 0.05
        k = int(rng.choice(len(opts), p=weights / weights.sum()))
        o = opts[k][0]
        if o == "none":
            out.append((x, y, w, h))
        elif o == "qt":
            for (dx, dy) in ((0, 0), (w // 2, 0), (0, h // 2), (w // 2, h // 2)):
                rec(x + dx, y + dy, w // 2, h // 2, depth + 1, True)
        elif o == "bth":
            rec(x, y, w, h // 2, depth + 1, False)
            rec(x, y + h // 2, w, h // 2, depth + 1, False)
        elif o == "btv":
            rec(x, y, w // 2, h, depth + 1, False)
            rec(x + w // 2, y, w // 2, h, depth + 1, False)
        elif o == "tth":
            q = h // 4
            rec(x, y, w, q, depth + 1, False)
            rec(x, y + q, w, 2 * q, depth + 1, False)
            rec(x, y + 3 * q, w, q, depth + 1, False)
        else:
            q = w // 4
            rec(x, y, q, h, depth + 1, False)
            rec(x + q, y, 2 * q, h, depth + 1, False)
            rec(x + 3 * q, y, q, h, depth + 1, False)

    rec(ctu_x, ctu_y, CTU, CTU, 0, True)
    return out


class FramePair:
    """Planes (Y, Cb, Cr) of the current and reference pictures plus blocks."""

    def __init__(self, cur, ref, blocks, ctu_mv):
        self.cur = cur            # list of 3 uint16 arrays
        self.ref = ref
        self.blocks = blocks      # int64 array (n, 8): x, y, w, h, mvx, mvy, mvpx, mvpy
        self.ctu_mv = ctu_mv      # per-CTU MV (1/16 pel)

    @property
    def width(self):
        return self.cur[0].shape[1]

    @property
    def height(self):
        return self.cur[0].shape[0]


def frame_pair_1080p(seed=2, width=1920, height=1080, mv_range_px=32, noise=4.0):
    """Smooth-field reference (Y, Cb, Cr); per-64x64-CTU motion uniform in
    +/-mv_range_px at 1/16 pel; current = reference displaced per CTU
    (8-tap luma, 4-tap chroma) plus N(0, noise^2); seeded QTMT blocks whose
    base MV is the CTU MV rounded to integer pel (also the predictor)."""
    cw, ch = width // 2, height // 2
    ref = [smooth_field(height, width, seed, 0), smooth_field(ch, cw, seed, 1), smooth_field(ch, cw, seed, 2)]
    nctx, ncty = (width + CTU - 1) // CTU, (height + CTU - 1) // CTU
    r = 16 * mv_range_px
    mv = randint(seed, 2 * nctx * ncty, -r, r, stream=3).reshape(ncty, nctx, 2)
    cur = [np.zeros((height, width), np.int64), np.zeros((ch, cw), np.int64), np.zeros((ch, cw), np.int64)]
    blocks = []
    for ty in range(ncty):
        for tx in range(nctx):
            mx, my = int(mv[ty, tx, 0]), int(mv[ty, tx, 1])
            x0, y0 = tx * CTU, ty * CTU
            w, h = min(CTU, width - x0), min(CTU, height - y0)
            cur[0][y0:y0 + h, x0:x0 + w] = _interp(ref[0], x0, y0, w, h, mx, my, 8, 4)
            for p in (1, 2):
                cur[p][y0 // 2:(y0 + h) // 2, x0 // 2:(x0 + w) // 2] = _interp(
                    ref[p], x0 // 2, y0 // 2, w // 2, h // 2, mx, my, 4, 5)
            bx, by = ((mx + 8) >> 4) << 4, ((my + 8) >> 4) << 4
            for (x, y, bw, bh) in qtmt_partition(x0, y0, width, height, seed * 1000003 + ty * nctx + tx):
                blocks.append((x, y, bw, bh, bx, by, bx, by))
    sig = [gaussian(seed, a.size, noise, stream=10 + i).reshape(a.shape) for i, a in enumerate(cur)]
    cur = [np.clip(np.rint(a + s), 0, 1023).astype(np.uint16) for a, s in zip(cur, sig)]
    ref = [a.astype(np.uint16) for a in ref]
    return FramePair(cur, ref, np.array(blocks, dtype=np.int64), mv)


def crop_pair(fp, width, height):
    """Blocks of fp inside the top-left width x height window (for small CPU checks)."""
    b = fp.blocks
    keep = (b[:, 0] + b[:, 2] <= width) & (b[:, 1] + b[:, 3] <= height)
    return b[keep]


# ---------------------------------------------------------------------------
# Forward transforms
# ---------------------------------------------------------------------------
# Relative frequency of each side length, VVC-like (small blocks dominate).
SIDE_WEIGHTS = {4: 0.30, 8: 0.30, 16: 0.22, 32: 0.13, 64: 0.05}


def residual_blocks(n=200000, seed=3, scale=12.0):
    """Block shapes over {4..64}^2 drawn with SIDE_WEIGHTS per side, mtsIdx 0
    (DCT-II) or 2..5 (DST-VII/DCT-VIII combinations) when both sides <= 32,
    Laplacian(scale) residuals clipped to [-1023, 1023] as int16.
    Returns (desc uint32 [lgW | lgH<<4 | mts<<8], offsets int64, residual int16)."""
    sides = np.array(sorted(SIDE_WEIGHTS), dtype=np.int64)
    p = np.array([SIDE_WEIGHTS[s] for s in sides])
    cdf = np.cumsum(p) / p.sum()
    u = uniform01(seed, 3 * n, stream=1).reshape(3, n)
    w = sides[np.searchsorted(cdf, u[0], side="right")]
    h = sides[np.searchsorted(cdf, u[1], side="right")]
    mts_choice = np.array([0, 2, 3, 4, 5], dtype=np.int64)
    mts = mts_choice[np.minimum((u[2] * 5).astype(np.int64), 4)]
    mts = np.where((w <= 32) & (h <= 32), mts, 0)
    lgw = np.log2(w).astype(np.int64)
    lgh = np.log2(h).astype(np.int64)
    sizes = w * h
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    total = int(sizes.sum())
    uu = uniform01(seed, total, stream=2) - 0.5
    lap = -scale * np.sign(uu) * np.log(np.maximum(1.0 - 2.0 * np.abs(uu), 1e-300))
    resid = np.clip(np.rint(lap), -1023, 1023).astype(np.int16)
    desc = (lgw | (lgh << 4) | (mts << 8)).astype(np.uint32)
    return desc, offsets, resid


# ---------------------------------------------------------------------------
# 4K layered scene: the full motion-search pipeline
# ---------------------------------------------------------------------------
PAN16 = (88, 36)                   # background pan per frame, 1/16 luma pel: (+5.5, +2.25) px
# 1/16-pel stage of the pipeline: the quarter-pel MV itself, then the rings of
# one and two sixteenths in the order L, R, T, B, TL, TR, BL, BR
CAND17_SIXTEENTH = [(0, 0)] + [(dx * r, dy * r) for r in (1, 2)
                               for (dx, dy) in ((-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1))]


def _layer(plane, tex, ox16, oy16, taps, fbits, margin):
    """Paint texture `tex` (with `margin` extra samples around it) whose
    top-left sits at (ox16, oy16) in 1/2^fbits sample units; clipped to the plane."""
    H, W = plane.shape
    th, tw = tex.shape[0] - 2 * margin, tex.shape[1] - 2 * margin
    ix, fx = ox16 >> fbits, ox16 & ((1 << fbits) - 1)
    iy, fy = oy16 >> fbits, oy16 & ((1 << fbits) - 1)
    x0, y0 = max(ix, 0), max(iy, 0)
    x1, y1 = min(ix + tw, W), min(iy + th, H)
    if x1 <= x0 or y1 <= y0:
        return
    # content at plane x is tex(x - ox): interpolate tex at offset -frac
    blk = _interp(tex, margin + (x0 - ix), margin + (y0 - iy), x1 - x0, y1 - y0, -fx, -fy, taps, fbits)
    plane[y0:y1, x0:x1] = blk


def layered_frames(width=3840, height=2160, seed=4, nobj=12, frames=3, max_mv=64, sides=(64, 512)):
    """Frames t = 0..frames-1 of a layered 10-bit 4:2:0 scene: a smooth background
    panned by (+5.5, +2.25) px per frame, and `nobj` seeded rectangles (sides in
    `sides`, own smooth textures, velocities uniform in +/-max_mv px per frame at
    quarter-pel precision) drawn in index order, so later ones occlude earlier
    ones.  Returns [(Y, Cb, Cr)] as uint16 planes, oldest first."""
    m = 16 + 2 * int(np.ceil(max(PAN16) * frames / 16.0))
    bgY = smooth_field(height + 2 * m, width + 2 * m, seed, 10)
    bgU = smooth_field(height // 2 + m, width // 2 + m, seed, 11)
    bgV = smooth_field(height // 2 + m, width // 2 + m, seed, 12)
    r = uniform01(seed, 8 * nobj, stream=13).reshape(nobj, 8)
    objs = []
    for k in range(nobj):
        w = int(sides[0] + r[k, 0] * (sides[1] - sides[0])) // 4 * 4
        h = int(sides[0] + r[k, 1] * (sides[1] - sides[0])) // 4 * 4
        px, py = int(r[k, 2] * (width - w)) // 2 * 2, int(r[k, 3] * (height - h)) // 2 * 2
        vx = int(np.rint((r[k, 4] * 2 - 1) * max_mv * 4)) * 4       # 1/16 pel per frame, quarter-pel steps
        vy = int(np.rint((r[k, 5] * 2 - 1) * max_mv * 4)) * 4
        tY = smooth_field(h + 16, w + 16, seed, 100 + k)
        tU = smooth_field(h // 2 + 8, w // 2 + 8, seed, 200 + k)
        tV = smooth_field(h // 2 + 8, w // 2 + 8, seed, 300 + k)
        objs.append((px, py, vx, vy, tY, tU, tV))
    out = []
    for t in range(frames):
        Y = _interp(bgY, m, m, width, height, -PAN16[0] * t, -PAN16[1] * t, 8, 4).astype(np.int64)
        U = _interp(bgU, m // 2, m // 2, width // 2, height // 2, -PAN16[0] * t, -PAN16[1] * t, 4, 5).astype(np.int64)
        V = _interp(bgV, m // 2, m // 2, width // 2, height // 2, -PAN16[0] * t, -PAN16[1] * t, 4, 5).astype(np.int64)
        for (px, py, vx, vy, tY, tU, tV) in objs:
            ox, oy = px * 16 + vx * t, py * 16 + vy * t            # 1/16 luma = 1/32 chroma units
            _layer(Y, tY, ox, oy, 8, 4, 8)
            _layer(U, tU, ox, oy, 4, 5, 4)
            _layer(V, tV, ox, oy, 4, 5, 4)
        out.append(tuple(np.ascontiguousarray(a.astype(np.uint16)) for a in (Y, U, V)))
    return out


def block_grid(width, height, bw=16, bh=16):
    """Raster grid of bw x bh blocks as int64 rows (x, y, w, h, 0, 0, 0, 0)."""
    ys, xs = np.mgrid[0:height // bh, 0:width // bw]
    b = np.zeros((xs.size, 8), np.int64)
    b[:, 0] = xs.ravel() * bw
    b[:, 1] = ys.ravel() * bh
    b[:, 2] = bw
    b[:, 3] = bh
    return b
