"""The oracle against golden vectors written by the reference C itself.

tests/golden/interp_golden.npz holds the outputs of the xin26x 10-bit build of
h266_core/source/h266_luma_mc.c and h266_chroma_mc.c (recipe oracle/Makefile,
generator tests/golden/make_golden.py) on seeded planes: every luma phase pair
under filterIndex 0/1/2, every chroma phase pair, every table width class
1..128, clipping extremes (flat 0/1023 and a 0/1023 checkerboard) and the
bi-prediction average / BCW weighting.

tests/golden/quant_golden.npz holds the outputs of the reference's own build of
h266_encoder/source/h266_quant_inv_quant.c (generator
tests/golden/make_quant_golden.py): every coefficient-group shape the encoder
passes, QPs 0..51 with the RDOQ / intra / inter rounding offsets, int16 extremes.

tests/golden/bili_golden.npz holds the outputs of the reference's own build of
h266_core/source/h266_bilinear_filter.c (generator tests/golden/make_bili_golden.py):
every DMVR bilinear phase pair, widths 1..128, flat and checkerboard extremes.
"""
import numpy as np

import golden_cases as gc


def _planes(golden):
    return golden["src_u16"], golden["src_s16"], golden["src2_s16"]


def test_golden_case_table_is_current(golden):
    assert np.array_equal(golden["cases"], gc.build_cases())
    planes = gc.make_planes()
    for a, b in zip(planes, _planes(golden)):
        assert np.array_equal(a, b)


def test_oracle_matches_reference_golden(oracle, golden):
    cases = golden["cases"]
    out_u, out_s, offs = gc.replay(gc.resolver(oracle, 1), cases, _planes(golden))
    assert np.array_equal(offs, golden["offsets"])
    bad = np.nonzero((out_u != golden["out_u16"]) | (out_s != golden["out_s16"]))[0]
    if bad.size:
        i = int(np.searchsorted(offs, bad[0], side="right") - 1)
        raise AssertionError(f"first mismatch in case {i}: {gc.KEYS[int(cases[i][0])]} "
                             f"params={cases[i].tolist()}")


def _case_outputs(golden, key, frac, fi, w=8, h=4):
    cases = golden["cases"]
    k = gc.KEYS.index(key)
    sel = np.nonzero((cases[:, 0] == k) & (cases[:, 1] == frac) & (cases[:, 4] == fi)
                     & (cases[:, 2] == w) & (cases[:, 3] == h))[0]
    return sel


def test_reference_hor_vet_copy_are_horvet_with_phase0(oracle, golden):
    """On the reference's own outputs: Hor(fx) = HorVet(fx, fy=0),
    Vet(fy) = HorVet(0, fy), Copy = HorVet(0, 0), for PIXEL and S16 outputs,
    every phase and filterIndex.  The HIP kernels compute every table slot
    with the HorVet form, so this identity is what makes that bit-exact.
    Checked by re-running the reference-equivalent HorVet at the same origin
    through the oracle (itself pinned by the test above)."""
    cases = golden["cases"]
    planes = _planes(golden)
    u = planes[0]
    import ctypes
    hv = {"u16": oracle.or_LumaInterpHorVet, "s16": oracle.or_LumaInterpHorVetU8S16}
    for f in hv.values():
        f.restype = None
        f.argtypes = [ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_void_p, ctypes.c_ssize_t,
                      ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    checked = 0
    for key, kout in (("LumaHor", "u16"), ("LumaVet", "u16"), ("LumaCopy", "u16"),
                      ("LumaHorS16", "s16"), ("LumaVetS16", "s16"), ("LumaCopyS16", "s16")):
        k = gc.KEYS.index(key)
        for i in np.nonzero(cases[:, 0] == k)[0]:
            frac, w, h, fi, ox, oy = (int(v) for v in cases[i][1:7])
            if key.startswith("LumaHor"):
                frac &= 15
            elif key.startswith("LumaVet"):
                frac &= 0xF0
            else:
                frac = 0
            dst = np.zeros((h, gc.OUT_STRIDE), np.uint16 if kout == "u16" else np.int16)
            hv[kout](ctypes.c_void_p(u.ctypes.data + 2 * (oy * gc.PLANE_W + ox)), gc.PLANE_W,
                     ctypes.c_void_p(dst.ctypes.data), gc.OUT_STRIDE, frac, w, h, fi)
            ref = golden["out_u16" if kout == "u16" else "out_s16"][golden["offsets"][i]:golden["offsets"][i + 1]]
            assert np.array_equal(dst[:, :w].reshape(-1), ref), (key, frac, w, h, fi)
            checked += 1
    assert checked > 300


def test_reference_horvet_phase_pairs_cover_all(golden):
    cases = golden["cases"]
    k = gc.KEYS.index("LumaHorVet")
    fr = set(int(c[1]) for c in cases if c[0] == k and c[4] == 0)
    assert fr == set(range(256))
    k = gc.KEYS.index("ChromaHorVet")
    fr = set(int(c[1]) for c in cases if c[0] == k)
    assert fr == set(range(1024))


def test_oracle_quant_matches_reference_golden(oracle):
    import os
    import quant_cases as qc
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "quant_golden.npz"))
    cases, total = qc.build_cases()
    assert np.array_equal(cases, g["cases"]) and np.array_equal(qc.make_inputs(total), g["t"])
    t = np.ascontiguousarray(g["t"])
    q, r, maps = qc.replay(qc.resolve(oracle, "or_QuantInvQuant"), cases, t)
    assert np.array_equal(q, g["q"]) and np.array_equal(r, g["r"])
    assert np.array_equal(maps, g["maps"])
    # the fixture exercises both clips and empty groups
    assert (g["q"] == 32767).any() or (g["q"] == -32768).any() or (np.abs(g["r"]) == 32767).any()
    assert (g["maps"] == 0).any() and (g["maps"] != 0).any()


def test_oracle_bilinear_matches_reference_golden(oracle):
    import os
    import bili_cases as bc
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "bili_golden.npz"))
    cases = bc.build_cases()
    assert np.array_equal(cases, g["cases"]) and np.array_equal(bc.make_plane(), g["plane"])
    out, offs = bc.replay(bc.resolve(oracle, "or_"), cases, np.ascontiguousarray(g["plane"]))
    assert np.array_equal(offs, g["offsets"])
    bad = np.nonzero(out != g["out"])[0]
    assert bad.size == 0, f"first mismatch in case {int(np.searchsorted(offs, bad[0], side='right') - 1)}"


def test_oracle_conventions_frozen(oracle):
    """Build-defined conventions of the closed reference functions, frozen
    (tests/golden/convention_cases.py): a change to any of them must be
    deliberate and regenerate the fixture."""
    import os
    import convention_cases as cc
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "convention_golden.npz"))
    got = cc.compute(oracle)
    assert set(got) == set(g.files)
    for k in g.files:
        assert np.array_equal(got[k], g[k]), k
