"""The C-ABI against the reference's own primitive table: every per-call
function of include/xg266.h is assigned, with no cast, to its
xin_func_struct field (h266_encoder/include/h266_func_struct.h:719-799) in a
C file that includes the reference's headers as h266_func_init.c does, built
10-bit (ENABLE_10BIT_ENCODER, video_api/video_common/include/xin_typedef.h:72-94)
with incompatible pointer assignments as errors.  The binding block of the
integration guide is compiled the same way.  Needs the reference sources, so it runs
where they are mounted and skips elsewhere.  CPU only."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("XG_REFERENCE_DIR", "/root/reference")
FUNC_INIT = os.path.join(REF, "h266_encoder", "source", "h266_func_init.c")

pytestmark = pytest.mark.skipif(not (os.path.isfile(FUNC_INIT) and shutil.which("gcc")),
                                reason="needs the reference sources and gcc")

INC = ["video_api/video_common/include", "video_api/video_encoder/include", "common/include", "h266_core/include",
       "h26x_core/include", "h266_encoder/include"]

# field expression (with i = width class, t = transform type) -> function
BY_NAME = """
    for (int i = 0; i < XIN_BLOCK_NUM; i++) {
        f->pfXinComputeSad[i] = Xg266ComputeSad;
        f->pfXinComputeSatd[i] = Xg266ComputeSatd;
        f->pfXinComputeDist[i] = Xg266ComputeSatd;
        f->pfXinComputeAvgSad[i] = Xg266ComputeAvgSad;
        f->pfXinComputeAvgSatd[i] = Xg266ComputeAvgSatd;
        f->pfXinComputeAvgDist[i] = Xg266ComputeAvgSatd;
        f->pfXinComputeSadx8[i] = Xg266ComputeSadx8;
        f->pfXinComputeSadx5[i] = Xg266ComputeSadx5;
        f->pfXinComputeSadx3[i] = Xg266ComputeSadx3;
        f->pfXinComputeAvgSadx8[i] = Xg266ComputeAvgSadx8;
        f->pfXinComputeSadS16[i] = Xg26xComputeSadS16;
        f->pfXinLumaInterp[0][0][i] = Xg266InterpCopy;
        f->pfXinLumaInterp[1][0][i] = Xg266LumaInterpHor;
        f->pfXinLumaInterp[0][1][i] = Xg266LumaInterpVet;
        f->pfXinLumaInterp[1][1][i] = Xg266LumaInterpHorVet;
        f->pfXinLumaInterpS16[0][0][i] = Xg266InterpCopyU8S16;
        f->pfXinLumaInterpS16[1][0][i] = Xg266LumaInterpHorU8S16;
        f->pfXinLumaInterpS16[0][1][i] = Xg266LumaInterpVetU8S16;
        f->pfXinLumaInterpS16[1][1][i] = Xg266LumaInterpHorVetU8S16;
        f->pfXinChromaInterp[0][0][i] = Xg266InterpCopy;
        f->pfXinChromaInterp[1][0][i] = Xg266ChromaInterpHor;
        f->pfXinChromaInterp[0][1][i] = Xg266ChromaInterpVet;
        f->pfXinChromaInterp[1][1][i] = Xg266ChromaInterpHorVet;
        f->pfXinChromaInterpS16[0][0][i] = Xg266InterpCopyU8S16;
        f->pfXinChromaInterpS16[1][0][i] = Xg266ChromaInterpHorU8S16;
        f->pfXinChromaInterpS16[0][1][i] = Xg266ChromaInterpVetU8S16;
        f->pfXinChromaInterpS16[1][1][i] = Xg266ChromaInterpHorVetU8S16;
        f->pfXinBiliInterp[0][0][i] = Xg266BiliInterpCopy;
        f->pfXinBiliInterp[1][0][i] = Xg266BiliInterpHor;
        f->pfXinBiliInterp[0][1][i] = Xg266BiliInterpVet;
        f->pfXinBiliInterp[1][1][i] = Xg266BiliInterpHorVet;
        f->pfXinInterpAvg[i] = Xg266InterpAvgS16U8;
        f->pfXinInterpWeight[i] = Xg266InterpWeightS16U8;
        f->pfXinQuantInvQuant[i] = Xg266QuantInvQuant;
        f->pfXinBlockSubForDct[i] = Xg26xBlockSubForDct;
        f->pfXinBlockAddForDct[i] = Xg26xBlockRecon;
        f->pfXinInv2dSkip[i] = Xg266ISkipWxH;
        for (int j = 0; j < XIN_BLOCK_NUM; j++) {
            f->pfXinFor2dDct2[i][j] = Xg26xFDct2WxH;
            f->pfXinInv2dDct2[i][j] = Xg266IDct2WxH;
        }
    }
    f->pfXinFor1dDct[XIN_DCT2][1] = Xg26xFDct2P2;
    f->pfXinFor1dDct[XIN_DCT2][2] = Xg26xFDct2P4;
    f->pfXinFor1dDct[XIN_DCT2][3] = Xg26xFDct2P8;
    f->pfXinFor1dDct[XIN_DCT2][4] = Xg26xFDct2P16;
    f->pfXinFor1dDct[XIN_DCT2][5] = Xg26xFDct2P32;
    f->pfXinFor1dDct[XIN_DCT2][6] = Xg26xFDct2P64;
    f->pfXinFor1dDct[XIN_DCT8][2] = Xg26xFDct8P4;
    f->pfXinFor1dDct[XIN_DCT8][3] = Xg26xFDct8P8;
    f->pfXinFor1dDct[XIN_DCT8][4] = Xg26xFDct8P16;
    f->pfXinFor1dDct[XIN_DCT8][5] = Xg26xFDct8P32;
    f->pfXinFor1dDct[XIN_DST7][2] = Xg26xFDst7P4;
    f->pfXinFor1dDct[XIN_DST7][3] = Xg26xFDst7P8;
    f->pfXinFor1dDct[XIN_DST7][4] = Xg26xFDst7P16;
    f->pfXinFor1dDct[XIN_DST7][5] = Xg26xFDst7P32;
    f->pfXinInv1dDct[XIN_DCT2][1] = Xg26xIDct2P2;
    f->pfXinInv1dDct[XIN_DCT2][2] = Xg26xIDct2P4;
    f->pfXinInv1dDct[XIN_DCT2][3] = Xg26xIDct2P8;
    f->pfXinInv1dDct[XIN_DCT2][4] = Xg26xIDct2P16;
    f->pfXinInv1dDct[XIN_DCT2][5] = Xg26xIDct2P32;
    f->pfXinInv1dDct[XIN_DCT2][6] = Xg26xIDct2P64;
    f->pfXinInv1dDct[XIN_DCT8][2] = Xg26xIDct8P4;
    f->pfXinInv1dDct[XIN_DCT8][3] = Xg26xIDct8P8;
    f->pfXinInv1dDct[XIN_DCT8][4] = Xg26xIDct8P16;
    f->pfXinInv1dDct[XIN_DCT8][5] = Xg26xIDct8P32;
    f->pfXinInv1dDct[XIN_DST7][2] = Xg26xIDst7P4;
    f->pfXinInv1dDct[XIN_DST7][3] = Xg26xIDst7P8;
    f->pfXinInv1dDct[XIN_DST7][4] = Xg26xIDst7P16;
    f->pfXinInv1dDct[XIN_DST7][5] = Xg26xIDst7P32;
    f->pfXinComputeHada8x8 = Xg266ComputeHada8x8;
    /* the S16-input second pass has no table slot; its one header declaration is the SSSE3 form */
    __typeof__(&Xin266LumaInterpVetS16U8_SSSE3) vetS16U8 = Xg266LumaInterpVetS16U8;
    (void)vetS16U8;
"""


def _prelude():
    """The reference's own include chain of its table translation unit (every
    #include of h266_func_init.c up to h266_func_struct.h)."""
    inc = []
    for line in open(FUNC_INIT):
        m = re.match(r'\s*#include\s+"([^"]+)"', line)
        if m:
            inc.append(m.group(1))
            if m.group(1) == "h266_func_struct.h":
                break
    assert "h266_func_struct.h" in inc
    return "".join(f'#include "{h}"\n' for h in inc) + '#include "xg266.h"\n'


def _compile(tmp_path, body, name):
    src = tmp_path / f"{name}.c"
    src.write_text(_prelude() + body)
    cmd = ["gcc", "-std=c11", "-DENABLE_10BIT_ENCODER", "-Wall", "-Wno-unused-function",
           "-Werror=incompatible-pointer-types", "-Werror=discarded-qualifiers", "-Werror=int-conversion",
           "-I", os.path.join(ROOT, "include")]
    for d in INC:
        cmd += ["-I", os.path.join(REF, d)]
    cmd += ["-c", str(src), "-o", str(tmp_path / f"{name}.o")]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_every_table_function_binds_without_casts(tmp_path):
    r = _compile(tmp_path, "void xg_bind_all(xin_func_struct *f)\n{\n" + BY_NAME + "}\n", "by_name")
    assert r.returncode == 0, r.stderr
    # every per-call table function the header declares is bound above
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "xg266.h")).read(), flags=re.S)
    declared = set(re.findall(r"\b(Xg26[6x][A-Za-z0-9]+)\s*\(", hdr))
    # entry points that are not table slots: the table helper, the ME and DMVR
    # entries (struct-based), the TU-level dispatchers, and the
    # S16-input second passes that h266_luma_mc.c / h266_chroma_mc.c define with
    # no header declaration (their argument lists are exercised by the golden
    # replay, which calls the reference's compiled functions the same way)
    not_table = {"Xg266FuncInit", "Xg26xMotionEstFull", "Xg26xMotionEstDia", "Xg266FullSubPelMe", "Xg266FDctWxH",
                 "Xg266IDctWxH", "Xg266IDctAdd", "Xg266DeriveMvFromDmvr", "Xg266LumaInterpVetS16S16",
                 "Xg266ChromaInterpVetS16U8", "Xg266ChromaInterpVetS16S16"}
    bound = set(re.findall(r"= (Xg26[6x][A-Za-z0-9]+);", BY_NAME))
    assert declared - not_table - bound == set()


def test_integration_binding_block_compiles(tmp_path):
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```c\n(#include \"xg266.h\"\n.*?)```", txt, re.S)
    assert m, "the integration guide has no binding block"
    block = m.group(1)
    assert not re.search(r"\((Xin|Xg)[A-Za-z0-9]+\)\s*g\.", block), "the binding block casts"
    r = _compile(tmp_path, block.replace('#include "xg266.h"\n', ""), "integration")
    assert r.returncode == 0, r.stderr


def test_a_wrong_slot_is_rejected(tmp_path):
    """The check has teeth: a function bound to a slot of another signature
    does not compile."""
    r = _compile(tmp_path, "void bad(xin_func_struct *f) { f->pfXinComputeSad[3] = Xg266ComputeSadx8; }\n", "bad")
    assert r.returncode != 0 and "incompatible" in r.stderr
