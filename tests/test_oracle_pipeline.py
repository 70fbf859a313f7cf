"""The oracle's motion-search pipeline recovers known motion: on a scene that
only pans by (+5.5, +2.25) px per frame, interior blocks find exactly that
motion at quarter pel against the previous frame and twice it against the
frame before; stages never increase their own cost.  CPU only."""
import numpy as np

from tests.pipeline_common import oracle_pipeline, scene


def test_pan_only_scene_recovers_the_pan(oracle):
    cur, ref0, ref1 = scene(w=192, h=96, nobj=0)
    out = oracle_pipeline(oracle, cur, ref0, ref1, rng=32)
    nbx, nby = 192 // 16, 96 // 16
    q = out["q"].reshape(nby, nbx, 2, 2)
    inner = q[1:-1, 1:-1]
    # content moves by +(5.5, 2.25) px per frame: the match in the previous frame is at -(5.5, 2.25) px
    assert np.all(inner[:, :, 0, 0] == -22) and np.all(inner[:, :, 0, 1] == -9)
    assert np.all(inner[:, :, 1, 0] == -44) and np.all(inner[:, :, 1, 1] == -18)
    # the refinement keeps the exact motion of the chosen reference, in 1/16 pel (ref1 lies an
    # integer 11 px away horizontally, so it predicts with less interpolation error and wins)
    ref = out["ref"].reshape(nby, nbx)[1:-1, 1:-1]
    bx, by = out["bx"].reshape(nby, nbx)[1:-1, 1:-1], out["by"].reshape(nby, nbx)[1:-1, 1:-1]
    assert np.array_equal(bx, np.where(ref == 0, -88, -176)) and np.array_equal(by, np.where(ref == 0, -36, -72))


def test_stage_costs_never_increase(oracle):
    cur, ref0, ref1 = scene()
    out = oracle_pipeline(oracle, cur, ref0, ref1)
    assert np.all(out["qCost"] <= out["intCost"])
    # the chosen reference is the one with the lower quarter-pel cost, ties to ref0
    qc = out["qCost"].reshape(-1, 2)
    assert np.array_equal(out["ref"], (qc[:, 1] < qc[:, 0]).astype(np.int32))
    # the final MV is one of the 1/16 candidates around 4x the chosen quarter-pel MV
    q = out["q"].reshape(-1, 2, 2)
    base = q[np.arange(len(q)), out["ref"]] * 4
    assert np.all(np.abs(out["bx"] - base[:, 0]) <= 2) and np.all(np.abs(out["by"] - base[:, 1]) <= 2)


def test_sampled_pipeline_equals_whole_frame(oracle):
    """or_MePipelineSample gives, for each sampled block, exactly what
    or_MePipeline gives for it (the sample's own and neighbours' stage-1
    searches are all a block's stages read)."""
    from tests.pipeline_common import oracle_pipeline_sample
    cur, ref0, ref1 = scene(w=192, h=96, nobj=3, max_mv=20)
    full = oracle_pipeline(oracle, cur, ref0, ref1, rng=24)
    nb = (192 // 16) * (96 // 16)
    sample = np.array([0, 5, 11, 12, 30, 47, 60, 71, 23, 12 * 5 + 6], np.int32)
    part = oracle_pipeline_sample(oracle, cur, ref0, ref1, sample, rng=24, threads=3)
    for k, width in (("full", 4), ("pred", 4), ("int", 4), ("q", 4), ("intCost", 2), ("qCost", 2)):
        assert np.array_equal(part[k].reshape(-1, width), full[k].reshape(nb, width)[sample]), k
    for k in ("ref", "bi", "bc", "bx", "by"):
        assert np.array_equal(part[k], full[k][sample]), k
    assert np.array_equal(part["cand"].reshape(len(sample), -1), full["cand"].reshape(nb, -1)[sample])
    for p in range(3):
        s = 1 if p == 0 else 2
        for b in sample.tolist():
            x, y = (b % 12) * 16 // s, (b // 12) * 16 // s
            assert np.array_equal(part["pred_pic"][p][y:y + 16 // s, x:x + 16 // s],
                                  full["pred_pic"][p][y:y + 16 // s, x:x + 16 // s])
