"""Independent numpy restatements of the reference's motion-search
compositions, used to pin the oracle's frame-level drivers.

Nothing here calls the oracle.  The parts are pinned on their own:
  * interp(): the 10-bit table dispatch of h266_core/source/h266_luma_mc.c and
    h266_chroma_mc.c ([!!fx][!!fy] -> Copy / Hor / Vet / HorVet), checked
    against the reference's own golden vectors (tests/golden/interp_golden.npz);
  * satd(): 2-D Walsh-Hadamard as a matrix product with scipy's Hadamard
    matrix (HM tiling and normalisation);
  * eg_bits(): lengths of explicitly built signed Exp-Golomb codewords.
The compositions are restated from the reference's source:
  * cu_window(): h266_encoder/source/h266_enc_init.c:1462-1465;
  * mc_split(): h266_encoder/source/h266_inter_prediction.c:332-342 (luma),
    :664-674 (chroma, 1/32 of a chroma sample);
  * refine_blocks(): cost (SATD << 8) + bits * lambda with strict < earliest
    (h266_sub_pel_refine.c:155-158), then the winner's MC;
  * full_sub_pel(): h266_encoder/source/h266_sub_pel_refine.c:81-212;
  * fill_me_cand(): h266_encoder/source/h266_analyze_inter_cu.c:133-293;
  * the integer searches and the pipeline follow the build-defined rules the
    closed Xin26xMotionEstFull / Xin26xMotionEstDia leave open, as written in
    the library header (raster full search, candidate list then L,R,T,B
    diamond steps, strict < everywhere).
Test infrastructure only."""
import numpy as np
import scipy.linalg

from xin26x_amd.workloads import CHROMA_FILTER, LUMA_FILTER

U32 = 0xFFFFFFFF
RING = [(-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1)]   # h266_sub_pel_refine.c:47-55


# ------------------------------------------------------------ primitives ---
def _se_codeword(v):
    k = 2 * v - 1 if v > 0 else -2 * v
    info = bin(k + 1)[3:]
    return "0" * len(info) + "1" + info


_EG = {}


def eg_bits(v):
    v = int(v)
    if v not in _EG:
        _EG[v] = len(_se_codeword(v))
    return _EG[v]


# The alternative half-pel luma filter (filterIndex 1 at phase 8; pinned by the
# reference's golden vectors with filterIndex 1).
LUMA_ALT_HPEL = (0, 3, 9, 20, 20, 9, 3, 0)


def interp(plane, x, y, w, h, frac, chroma, fi=0):
    """w x h prediction whose integer position is plane[y, x] (the reference's
    src pointer), 10-bit, filterIndex 0 (or 1: the alternative half-pel luma
    filter at phase 8).  `plane` must hold the taps' reach."""
    taps, shift, tab = (4, 5, CHROMA_FILTER) if chroma else (8, 4, LUMA_FILTER)
    if fi == 1 and not chroma:
        tab = [LUMA_ALT_HPEL if k == 8 else tab[k] for k in range(16)]
    mask = (1 << shift) - 1
    fx, fy = frac & mask, frac >> shift
    o = taps // 2 - 1                       # 3 for luma, 1 for chroma
    s = plane.astype(np.int64)
    if fx == 0 and fy == 0:                 # Copy
        return s[y:y + h, x:x + w].copy()
    if fy == 0:                             # Hor: clip((sum + 32) >> 6)
        acc = sum(int(tab[fx][k]) * s[y:y + h, x - o + k:x - o + k + w] for k in range(taps))
        return np.clip((acc + 32) >> 6, 0, 1023)
    if fx == 0:                             # Vet
        acc = sum(int(tab[fy][k]) * s[y - o + k:y - o + k + h, x:x + w] for k in range(taps))
        return np.clip((acc + 32) >> 6, 0, 1023)
    # HorVet: first pass (sum - (8192 << 2)) >> 2 over h + taps - 1 rows,
    # second pass (sum + (1 << 9) + (8192 << 6)) >> 10, clipped
    rows = s[y - o:y - o + h + taps - 1]
    t = sum(int(tab[fx][k]) * rows[:, x - o + k:x - o + k + w] for k in range(taps))
    t = (t - (8192 << 2)) >> 2
    acc = sum(int(tab[fy][k]) * t[k:k + h] for k in range(taps))
    return np.clip((acc + (1 << 9) + (8192 << 6)) >> 10, 0, 1023)


_HAD = {n: scipy.linalg.hadamard(n).astype(np.int64) for n in (2, 4, 8)}


def satd(a, b):
    """HM SATD: 8x8 tiles when both sides are multiples of 8, else 4x4, else
    2x2; per tile sum |H d H^T|, normalised (s+2)>>2, (s+1)>>1, none."""
    d = a.astype(np.int64) - b.astype(np.int64)
    h, w = d.shape
    n = 8 if (w % 8 == 0 and h % 8 == 0) else 4 if (w % 4 == 0 and h % 4 == 0) else 2
    H = _HAD[n]
    t = d.reshape(h // n, n, w // n, n).transpose(0, 2, 1, 3)
    s = np.abs(H @ t @ H.T).sum(axis=(2, 3))
    if n == 8:
        s = (s + 2) >> 2
    elif n == 4:
        s = (s + 1) >> 1
    return int(s.sum())


def sad(a, b):
    return int(np.abs(a.astype(np.int64) - b.astype(np.int64)).sum())


def int16(v):
    return ((int(v) + 32768) & 0xFFFF) - 32768


def cu_window(x, y, w, h, fw, fh, ctu):
    """(minX, minY, maxX, maxY) in integer pel, h266_enc_init.c:1462-1465 with
    XIN_PADDING_OFFSET_X/Y = 16 (common/include/h26x_definition.h:103-104)."""
    return (int16(-x - ctu - (16 - 8)), int16(-y - ctu - (16 - 8)),
            int16(fw + (ctu - w + 16 - 8) - x), int16(fh + (ctu - h + 16 - 8) - y))


def mc_split(x, y, mvx, mvy, chroma):
    """(integer x, integer y, frac) of a 1/16-pel luma MV at luma (x, y):
    refOff = (cuPel << 4) + mv; luma frac = refOff & 15 and int = refOff >> 4;
    chroma frac = refOff & 31 and int = refOff >> 5 (1/32 chroma sample)."""
    rx, ry = (x << 4) + mvx, (y << 4) + mvy
    if chroma:
        return rx >> 5, ry >> 5, ((ry & 31) << 5) | (rx & 31)
    return rx >> 4, ry >> 4, ((ry & 15) << 4) | (rx & 15)


class Pic:
    """Unpadded planes plus edge-padded copies; padded(p)[pad_p + y, pad_p + x]
    is sample (x, y)."""

    def __init__(self, planes, pad):
        self.planes = planes
        self.pads = [pad, pad // 2, pad // 2]
        self.padded = [np.pad(p.astype(np.int64), pp, mode="edge") for p, pp in zip(planes, self.pads)]

    def pred(self, p, x, y, w, h, frac):
        pp = self.pads[p]
        return interp(self.padded[p], x + pp, y + pp, w, h, frac, p > 0)


# ----------------------------------------------------------- refinement ---
def refine_blocks(cur, ref, blocks, offs, lam, ctu, use_chroma=True, pred=None):
    """Per block: every candidate MV base + offset (1/16 pel) clipped to the CU
    window, luma (+ chroma) prediction, SATD, cost (SATD << 8) + (EG(dx) +
    EG(dy)) * lam against the block's predictor, strict < keeps the earliest;
    the winner's luma and chroma prediction is written into `pred` (three
    planes).  Returns (costs [n, ncand], best index, best cost, best mvx, mvy)."""
    fh, fw = cur.planes[0].shape
    n = len(blocks)
    costs = np.zeros((n, len(offs)), np.int64)
    bi = np.zeros(n, np.int64)
    bc = np.zeros(n, np.int64)
    bmv = np.zeros((n, 2), np.int64)
    for i, (x, y, w, h, mvx0, mvy0, mvpx, mvpy) in enumerate(np.asarray(blocks, np.int64).tolist()):
        mnx, mny, mxx, mxy = cu_window(x, y, w, h, fw, fh, ctu)
        best, bidx, bm = None, 0, (0, 0)
        for c, (dx, dy) in enumerate(offs):
            mvx = min(max(mvx0 + int(dx), mnx * 16), mxx * 16)
            mvy = min(max(mvy0 + int(dy), mny * 16), mxy * 16)
            ix, iy, fr = mc_split(x, y, mvx, mvy, False)
            s = satd(cur.planes[0][y:y + h, x:x + w], ref.pred(0, ix, iy, w, h, fr))
            if use_chroma:
                cx, cy, frc = mc_split(x, y, mvx, mvy, True)
                for p in (1, 2):
                    s += satd(cur.planes[p][y // 2:(y + h) // 2, x // 2:(x + w) // 2],
                              ref.pred(p, cx, cy, w // 2, h // 2, frc))
            cost = ((s << 8) + (eg_bits(mvx - mvpx) + eg_bits(mvy - mvpy)) * lam) & U32
            costs[i, c] = cost
            if best is None or cost < best:
                best, bidx, bm = cost, c, (mvx, mvy)
        bi[i], bc[i], bmv[i] = bidx, best, bm
        if pred is not None:
            ix, iy, fr = mc_split(x, y, bm[0], bm[1], False)
            pred[0][y:y + h, x:x + w] = ref.pred(0, ix, iy, w, h, fr)
            cx, cy, frc = mc_split(x, y, bm[0], bm[1], True)
            for p in (1, 2):
                pred[p][y // 2:(y + h) // 2, x // 2:(x + w) // 2] = ref.pred(p, cx, cy, w // 2, h // 2, frc)
    return costs, bi, bc, bmv[:, 0], bmv[:, 1]


# ------------------------------------------------------------ sub-pel ---
def full_sub_pel(cur_blk, ref, x, y, mv, best_cost, mvp, lam, win):
    """Xin266FullSubPelMe (h266_sub_pel_refine.c:81-212) for the block whose
    current samples are cur_blk at luma (x, y): nothing when the integer part
    of the quarter-pel best MV is outside the window (:127-130); otherwise the
    half-pel ring (+/-2 quarter pel, order L,R,T,B,TL,TR,BL,BR) around the
    best, then the quarter-pel ring (+/-1) around the best after the half-pel
    ring; each candidate's SAD against the 8-tap prediction at
    frac = ((4 mvY) & 15) << 4 | ((4 mvX) & 15); cost (SAD << 8) +
    bits(mv - mvp) * lam; strict < (:158, :199).  Returns (mv, cost)."""
    h, w = cur_blk.shape
    mnx, mny, mxx, mxy = win
    mvx, mvy = mv
    if not (mnx <= (mvx >> 2) <= mxx and mny <= (mvy >> 2) <= mxy):
        return (mvx, mvy), best_cost
    best, bmx, bmy = best_cost, mvx, mvy
    for step in (2, 1):
        cx, cy = bmx, bmy
        for (dx, dy) in RING:
            tx, ty = cx + dx * step, cy + dy * step
            q16x, q16y = tx * 4, ty * 4
            p = ref.pred(0, x + (q16x >> 4), y + (q16y >> 4), w, h, ((q16y & 15) << 4) | (q16x & 15))
            cost = ((sad(cur_blk, p) << 8) + (eg_bits(tx - mvp[0]) + eg_bits(ty - mvp[1])) * lam) & U32
            if cost < best:
                best, bmx, bmy = cost, tx, ty
    return (bmx, bmy), best


# ------------------------------------------------------- integer search ---
def _int_cost(cur_blk, ref, x, y, mx, my, mvp, lam):
    h, w = cur_blk.shape
    pp = ref.pads[0]
    r = ref.padded[0][pp + y + my:pp + y + my + h, pp + x + mx:pp + x + mx + w]
    s = sad(cur_blk, r)
    return ((s << 8) + (eg_bits(4 * mx - mvp[0]) + eg_bits(4 * my - mvp[1])) * lam) & U32, s


def full_search(cur_blk, ref, x, y, win, mvp, rng, lam, clip_centre=False):
    """Raster full search (dy outer, dx inner, ascending) over +/-rng around
    the rounded predictor ((v + 2) >> 2, quarter pel in) inside the window;
    strict <.  clip_centre: the rounded centre is first clipped into the
    window (the serving search).  Returns ((mx, my) integer pel, cost, sad)."""
    h, w = cur_blk.shape
    mnx, mny, mxx, mxy = win
    cx, cy = (mvp[0] + 2) >> 2, (mvp[1] + 2) >> 2
    if clip_centre:
        cx, cy = min(max(cx, mnx), mxx), min(max(cy, mny), mxy)
    x0, x1 = max(cx - rng, mnx), min(cx + rng, mxx)
    y0, y1 = max(cy - rng, mny), min(cy + rng, mxy)
    pp = ref.pads[0]
    P = ref.padded[0]
    # every SAD of the window at once
    win_s = P[pp + y + y0:pp + y + y1 + h, pp + x + x0:pp + x + x1 + w]
    view = np.lib.stride_tricks.sliding_window_view(win_s, (h, w))
    sads = np.abs(view - cur_blk.astype(np.int64)).sum(axis=(2, 3))
    best = None
    for j, my in enumerate(range(y0, y1 + 1)):
        by = eg_bits(4 * my - mvp[1])
        for i, mx in enumerate(range(x0, x1 + 1)):
            s = int(sads[j, i])
            c = ((s << 8) + (eg_bits(4 * mx - mvp[0]) + by) * lam) & U32
            if best is None or c < best[1]:
                best = ((mx, my), c, s)
    return best


def diamond(cur_blk, ref, x, y, win, mvp, rng, lam, cands):
    """Start candidates in list order (each clipped into the window), then
    small-diamond steps L,R,T,B around the best, inside the window and the
    +/-rng box around the rounded predictor, until no neighbour is strictly
    better or 2 * rng steps were taken.  Returns ((mx, my), cost, sad)."""
    mnx, mny, mxx, mxy = win
    cx, cy = (mvp[0] + 2) >> 2, (mvp[1] + 2) >> 2
    best = None
    for (mx, my) in cands:
        mx, my = min(max(mx, mnx), mxx), min(max(my, mny), mxy)
        c, s = _int_cost(cur_blk, ref, x, y, mx, my, mvp, lam)
        if best is None or c < best[1]:
            best = ((mx, my), c, s)
    for _ in range(2 * rng):
        (bx, by), bc, _s = best
        nb = best
        for (dx, dy) in RING[:4]:
            mx, my = bx + dx, by + dy
            if not (mnx <= mx <= mxx and mny <= my <= mxy):
                continue
            if not (cx - rng <= mx <= cx + rng and cy - rng <= my <= cy + rng):
                continue
            c, s = _int_cost(cur_blk, ref, x, y, mx, my, mvp, lam)
            if c < nb[1]:
                nb = ((mx, my), c, s)
        if nb[0] == best[0]:
            break
        best = nb
    return best


def fill_me_cand(pred, extra):
    """Xin266FillMeCand (h266_analyze_inter_cu.c:133-293): (0, 0) first, then
    the predictor and the extra quarter-pel MVs rounded (v + 2) >> 2
    (:271-272), each appended unless already present (:265-291), at most 12."""
    out = [(0, 0)]
    for (vx, vy) in [pred] + list(extra):
        c = ((vx + 2) >> 2, (vy + 2) >> 2)
        if c not in out and len(out) < 12:
            out.append(c)
    return out


def median3(a, b, c):
    return sorted((a, b, c))[1]


def pipeline_cands(full, nbx, b, r):
    """Predictor (quarter pel) and start candidates (integer pel) of block b,
    reference r, from the stage-1 MVs full[block, ref] (integer pel): the
    component-wise median of the left, top and top-right MVs (missing ones
    count as (0, 0)), then fill_me_cand(predictor; own, left, top, top-right)."""
    bx, by = b % nbx, b // nbx
    nbr = [b, b - 1 if bx > 0 else -1, b - nbx if by > 0 else -1,
           b - nbx + 1 if (by > 0 and bx + 1 < nbx) else -1]
    q = [(0, 0) if k < 0 else (4 * int(full[k, r, 0]), 4 * int(full[k, r, 1])) for k in nbr]
    p = (median3(q[1][0], q[2][0], q[3][0]), median3(q[1][1], q[2][1], q[3][1]))
    return p, fill_me_cand(p, q)


def me_pipeline(cur, refs, bw, bh, rng, lam, ctu, offs, lam_refine):
    """The two-reference pipeline over a raster grid of bw x bh blocks:
      1. full search +/-rng around (0, 0) per block and reference;
      2. predictor = component-wise median of the left, top and top-right
         blocks' stage-1 MVs (missing ones count as (0, 0)), quarter pel;
         candidates = fill_me_cand(predictor; own, left, top, top-right
         stage-1 MVs); diamond search;
      3. half- then quarter-pel rings from the diamond result;
      4. the reference with the lower quarter-pel cost (ties: reference 0);
      5. refine_blocks() of the chosen reference with base MV 4 * quarter-pel
         MV and predictor 4 * that reference's predictor.
    Returns a dict of per-stage arrays."""
    fh, fw = cur.planes[0].shape
    nbx, nby = fw // bw, fh // bh
    nb = nbx * nby
    full = np.zeros((nb, 2, 2), np.int64)
    for r in range(2):
        for b in range(nb):
            x, y = (b % nbx) * bw, (b // nbx) * bh
            win = cu_window(x, y, bw, bh, fw, fh, ctu)
            full[b, r] = full_search(cur.planes[0][y:y + bh, x:x + bw], refs[r], x, y, win, (0, 0), rng, lam)[0]
    pmv = np.zeros((nb, 2, 2), np.int64)
    imv = np.zeros((nb, 2, 2), np.int64)
    icost = np.zeros((nb, 2), np.int64)
    qmv = np.zeros((nb, 2, 2), np.int64)
    qcost = np.zeros((nb, 2), np.int64)
    for r in range(2):
        for b in range(nb):
            x, y = (b % nbx) * bw, (b // nbx) * bh
            p, cands = pipeline_cands(full, nbx, b, r)
            pmv[b, r] = p
            win = cu_window(x, y, bw, bh, fw, fh, ctu)
            blk = cur.planes[0][y:y + bh, x:x + bw]
            (mx, my), c, _s = diamond(blk, refs[r], x, y, win, p, rng, lam, cands)
            imv[b, r] = (mx, my)
            icost[b, r] = c
            (qx, qy), qc = full_sub_pel(blk, refs[r], x, y, (4 * mx, 4 * my), c, p, lam, win)
            qmv[b, r] = (qx, qy)
            qcost[b, r] = qc
    ref_idx = (qcost[:, 1] < qcost[:, 0]).astype(np.int64)
    pred = [np.zeros_like(p) for p in cur.planes]
    cand = np.zeros((nb, len(offs)), np.int64)
    best = np.zeros((nb, 4), np.int64)
    for b in range(nb):
        r = int(ref_idx[b])
        x, y = (b % nbx) * bw, (b // nbx) * bh
        blk = [(x, y, bw, bh, 4 * qmv[b, r, 0], 4 * qmv[b, r, 1], 4 * pmv[b, r, 0], 4 * pmv[b, r, 1])]
        c, bi, bc, bmx, bmy = refine_blocks(cur, refs[r], blk, offs, lam_refine, ctu, True, pred)
        cand[b] = c[0]
        best[b] = (bi[0], bc[0], bmx[0], bmy[0])
    return dict(full=full, pred=pmv, int=imv, icost=icost, q=qmv, qcost=qcost, ref=ref_idx, cand=cand, best=best,
                pred_planes=pred)


# ------------------------------------------------------------------- AMVP ---
AMVR_PREC = (4, 2, 0, 3)      # amvrPrecision by imvIdx, h266_entropy_write.c:120-123


def change_prec(v, in_prec, out_prec):
    """One MV component from precision in_prec to out_prec (log2 units per 4
    pel): exact when the precision rises, else the nearest value with ties
    toward zero (the rounding of Xin266ChangeMv32Prec,
    h266_get_neighbour_mv.c:212-243), written with exact fractions."""
    from fractions import Fraction
    s = out_prec - in_prec
    if s >= 0:
        return v * (1 << s)
    q = Fraction(v, 1 << -s)
    lo = q.numerator // q.denominator
    hi = lo + (q != lo)
    if q - lo < hi - q:
        return lo
    if q - lo > hi - q:
        return hi
    return lo if abs(lo) < abs(hi) else hi


def amvp_list(field, fw, fh, x, y, w, h):
    """AMVP candidates of a block from an MV field [uy, ux, 2] of 4x4 units
    (every unit inside the picture an available neighbour), restated from
    Xin266FillAmvpCand (h266_get_neighbour_mv.c:2451-2560): below-left else
    left; above-right else above else above-left; a duplicate pair kept once;
    a zero MV appended to an empty list or to a single non-zero entry; the
    second entry kept only when it differs from the first."""
    def at(px, py):
        if 0 <= px < fw and 0 <= py < fh:
            return (int(field[py >> 2, px >> 2, 0]), int(field[py >> 2, px >> 2, 1]))
        return None
    out = []
    for group in (((x - 1, y + h), (x - 1, y + h - 1)),
                  ((x + w, y - 1), (x + w - 1, y - 1), (x - 1, y - 1))):
        for px, py in group:
            mv = at(px, py)
            if mv is not None:
                out.append(mv)
                break
    if len(out) == 2 and out[0] == out[1]:
        out = out[:1]
    if not out or (len(out) == 1 and out[0] != (0, 0)):
        out.append((0, 0))
    return out if len(out) >= 2 and out[0] != out[1] else out[:1]


def select_mvp(cur, ref, pad, x, y, w, h, cands, imv, ctu):
    """Xin266SelectMvp (h266_analyze_inter_cu.c:60-131) of one block: each
    candidate rounded to the AMVR precision and back, its luma prediction
    (MV clipped to the CU window, h266_inter_prediction.c:332-342; the
    alternative half-pel filter for imvIdx 3) compared by SAD, the first
    strictly lowest kept.  ref: the reference plane padded by `pad`.
    Returns (index, sad, rounded MV, MV in quarter pel)."""
    fh, fw = cur.shape
    mnx, mny, mxx, mxy = cu_window(x, y, w, h, fw, fh, ctu)
    prec = AMVR_PREC[imv]
    best = None
    for i, (cx, cy) in enumerate(cands):
        rx = change_prec(change_prec(cx, 6, prec), prec, 6)
        ry = change_prec(change_prec(cy, 6, prec), prec, 6)
        ox = 16 * x + min(max(rx, 16 * mnx), 16 * mxx)
        oy = 16 * y + min(max(ry, 16 * mny), 16 * mxy)
        p = interp(ref, (ox >> 4) + pad, (oy >> 4) + pad, w, h, ((oy & 15) << 4) | (ox & 15), False,
                   1 if imv == 3 else 0)
        s = sad(cur[y:y + h, x:x + w], p)
        if best is None or s < best[1]:
            best = (i, s, (rx, ry))
    i, s, (rx, ry) = best
    return i, s, (rx, ry), (change_prec(rx, 6, 4), change_prec(ry, 6, 4))


# ------------------------------------------------------------------- MCTF ---
def mctf_subsample(src, pad):
    """Rounded 2x2 means of src (h/2 x w/2), edge-padded by `pad` samples."""
    h, w = src.shape
    s = src[:h // 2 * 2, :w // 2 * 2].astype(np.int64)
    m = (s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2] + 2) >> 2
    return np.pad(m, pad, mode="edge").astype(np.uint16)


def _mctf_ssd(cur, refp, pad, bs, x, y, mx, my):
    """SSD of the block at (x, y) against the 8-tap prediction at (mx, my) in
    1/16 pel, or None when the 8-tap reach leaves the padding."""
    h, w = cur.shape
    px, py = x + (mx >> 4), y + (my >> 4)
    if px - 3 < -pad or py - 3 < -pad or px + bs + 4 > w - 1 + pad or py + bs + 4 > h - 1 + pad:
        return None
    p = interp(refp, px + pad, py + pad, bs, bs, ((my & 15) << 4) | (mx & 15), False)
    d = cur[y:y + bs, x:x + bs].astype(np.int64) - p
    return int((d * d).sum())


def mctf_level(cur, refp, pad, bs, rng, lower, factor, frac):
    """The library's MCTF search at one level, restated: per block, the 3x3
    lower-level MVs around (bx / 2, by / 2) times factor and zero, then the
    integer positions +/-rng around the best, then (frac) 7x7 grids at steps
    of 4 and 1 (1/16 pel) around the best; first strictly lowest SSD.
    Returns (mv [nby, nbx, 2], err [nby, nbx])."""
    h, w = cur.shape
    nbx, nby = w // bs, h // bs
    mv = np.zeros((nby, nbx, 2), np.int64)
    err = np.zeros((nby, nbx), np.int64)

    for by in range(nby):
        for bx in range(nbx):
            x, y = bx * bs, by * bs
            best, bm = None, (0, 0)

            def test(mx, my):
                nonlocal best, bm
                e = _mctf_ssd(cur, refp, pad, bs, x, y, mx, my)
                if e is not None and (best is None or e < best):
                    best, bm = e, (mx, my)

            if lower is not None:
                lh, lw = lower.shape[:2]
                for py in (-1, 0, 1):
                    for px in (-1, 0, 1):
                        tx, ty = bx // 2 + px, by // 2 + py
                        if 0 <= tx < lw and 0 <= ty < lh:
                            test(int(lower[ty, tx, 0]) * factor, int(lower[ty, tx, 1]) * factor)
                test(0, 0)
            cx, cy = int(bm[0] / 16), int(bm[1] / 16)          # C division: toward zero
            for y2 in range(cy - rng, cy + rng + 1):
                for x2 in range(cx - rng, cx + rng + 1):
                    test(16 * x2, 16 * y2)
            if frac:
                for step in (4, 1):
                    ox, oy = bm
                    for y2 in range(-3 * step, 3 * step + 1, step):
                        for x2 in range(-3 * step, 3 * step + 1, step):
                            test(ox + x2, oy + y2)
            mv[by, bx] = bm
            err[by, bx] = 0xFFFFFFFF if best is None else best
    return mv, err
