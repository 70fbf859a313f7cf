"""ctypes loader for the CPU oracle (oracle/liboracle.so): the checker.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

P = ctypes.c_void_p
SS = ctypes.c_ssize_t
I32 = ctypes.c_int32
U32 = ctypes.c_uint32


class OrBlock(ctypes.Structure):
    _fields_ = [("x", I32), ("y", I32), ("w", I32), ("h", I32),
                ("mvx", I32), ("mvy", I32), ("mvpx", I32), ("mvpy", I32)]


class OrRefinePrm(ctypes.Structure):
    _fields_ = [("ncand", I32), ("offx", P), ("offy", P), ("lambda_", U32),
                ("ctuSize", I32), ("useChroma", I32)]


class OrMe(ctypes.Structure):
    _fields_ = [("width", U32), ("height", U32),
                ("minMvX", ctypes.c_int16), ("minMvY", ctypes.c_int16),
                ("maxMvX", ctypes.c_int16), ("maxMvY", ctypes.c_int16),
                ("candX", ctypes.c_int16 * 12), ("candY", ctypes.c_int16 * 12),
                ("candNum", U32),
                ("bestMvX", ctypes.c_int16), ("bestMvY", ctypes.c_int16),
                ("predMvX", ctypes.c_int16), ("predMvY", ctypes.c_int16),
                ("bestCost", U32), ("bestSad", U32),
                ("searchWidth", I32), ("searchHeight", I32),
                ("sadLambda", U32),
                ("input", P), ("inputStride", SS),
                ("ref", P), ("refStride", SS)]


class OrDmvr(ctypes.Structure):
    _fields_ = [("ref0", P), ("stride0", SS), ("ref1", P), ("stride1", SS),
                ("cuPelX", I32), ("cuPelY", I32), ("width", U32), ("height", U32),
                ("mv0x", I32), ("mv0y", I32), ("mv1x", I32), ("mv1y", I32),
                ("minMvX", ctypes.c_int16), ("minMvY", ctypes.c_int16),
                ("maxMvX", ctypes.c_int16), ("maxMvY", ctypes.c_int16)]


def load_oracle(path=None):
    """The oracle library with every prototype set; `path` loads another
    build of the same source (the mutation tests build planted variants)."""
    if path is None and not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    lib = ctypes.CDLL(path or ORACLE_SO)
    sig = {
        "or_ComputeSad": (U32, [P, SS, P, SS, U32, U32]),
        "or_ComputeSatd": (U32, [P, SS, P, SS, U32, U32]),
        "or_ComputeHada8x8": (U32, [P, SS]),
        "or_ComputeSadx8": (None, [P, SS, P, SS, U32, U32, P]),
        "or_ComputeAvgSad": (U32, [P, SS, P, P, SS, U32, U32]),
        "or_ComputeAvgSatd": (U32, [P, SS, P, P, SS, U32, U32]),
        "or_ComputeAvgSadx8": (None, [P, SS, P, P, SS, U32, U32, P]),
        "or_ComputeSadx5": (None, [P, SS, P, P, P, P, P, SS, U32, U32, P]),
        "or_ComputeSadx3": (None, [P, SS, P, P, P, SS, U32, U32, P]),
        "or_ExpGolombBits": (U32, [I32]),
        "or_MvdBits": (U32, [I32, I32]),
        "or_SadLambda": (U32, [I32, I32]),
        "or_MvRoundShift": (I32, [I32, I32]),
        "or_ChangeMvPrec": (I32, [I32, I32, I32]),
        "or_MctfSubsample": (None, [P, SS, I32, I32, P, SS, I32]),
        "or_MctfMeLevel": (None, [P, SS, P, SS, I32, I32, I32, I32, I32, P, I32, I32, I32, I32, I32, I32, P, P]),
        "or_AmvpFromField": (I32, [P, SS, I32, I32, I32, I32, I32, I32, P]),
        "or_SelectMvp": (I32, [P, SS, P, SS, I32, I32, I32, I32, I32, I32, I32, P, I32, I32, P, P, P]),
        "or_MotionEstFull": (None, [ctypes.POINTER(OrMe)]),
        "or_MotionEstDia": (None, [ctypes.POINTER(OrMe)]),
        "or_FullSubPelMe": (None, [ctypes.POINTER(OrMe)]),
        "or_PipelineCands": (None, [P, I32, I32, I32, P, P, P, P, P]),
        "or_FillMeCand": (None, [ctypes.c_int16, ctypes.c_int16, P, P, U32, P, P, P]),
        "or_TrMatrix": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P]),
        "or_FDct1d": (None, [ctypes.c_int, ctypes.c_int, P, SS, P, SS, I32, I32]),
        "or_FDct2WxH": (None, [P, SS, P, SS, U32, U32, P, I32, I32]),
        "or_FDctWxH": (None, [P, SS, P, SS, U32, U32, U32, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "or_FDctBatch": (None, [P, P, P, I32, I32, P]),
        "or_IDct1d": (None, [ctypes.c_int, ctypes.c_int, P, SS, P, SS, I32, I32, I32]),
        "or_IDct2WxH": (None, [P, SS, P, SS, U32, U32, P, I32, I32]),
        "or_ISkipWxH": (None, [P, SS, P, SS, U32, U32]),
        "or_IDctWxH": (ctypes.c_int, [P, SS, P, SS, U32, U32, U32, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "or_BlockRecon": (None, [P, SS, P, SS, P, SS, U32, U32, ctypes.c_int]),
        "or_ReconBatch": (None, [P, P, P, P, I32, I32, P]),
        "or_BiliInterpCopy": (None, [P, SS, P, SS, I32, U32, U32]),
        "or_BiliInterpHor": (None, [P, SS, P, SS, I32, U32, U32]),
        "or_BiliInterpVet": (None, [P, SS, P, SS, I32, U32, U32]),
        "or_BiliInterpHorVet": (None, [P, SS, P, SS, I32, U32, U32]),
        "or_ComputeSadS16": (U32, [P, SS, P, SS, U32, U32, U32]),
        "or_DeriveMvFromDmvr": (None, [ctypes.POINTER(OrDmvr), P]),
        "or_DmvrSubPel": (None, [ctypes.c_int, P, P]),
        "or_Downscale2x2": (None, [P, SS, P, SS, U32, U32]),
        "or_IDctAdd": (ctypes.c_int, [P, SS, P, SS, P, SS, U32, U32, U32, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]),
        "or_QuantInvQuant": (None, [P, P, P, SS, U32, U32, U32, U32, I32, I32, I32, I32, I32, I32, P]),
        "or_QuantParams": (None, [U32, P]),
        "or_QuantBatch": (None, [P, P, P, I32, I32, P, P, P]),
        "or_MePipeline": (None, [P, P, P, SS, SS, P, P, P, P, P, P, SS, SS, I32, I32, I32, I32, I32, U32, I32,
                                 ctypes.POINTER(OrRefinePrm)] + [P] * 15 + [SS, SS]),
        "or_MePipelineSample": (None, [P, P, P, SS, SS, P, P, P, P, P, P, SS, SS, I32, I32, I32, I32, I32, U32, I32,
                                       ctypes.POINTER(OrRefinePrm), P, I32] + [P] * 15 + [SS, SS]),
        "or_BlockSubForDct": (None, [P, SS, P, SS, P, SS, U32, U32]),
        "or_MtsTypes": (None, [U32, P, P]),
        "or_LumaInterpolation": (None, [P, SS, P, SS, I32, U32, U32, U32]),
        "or_LumaInterpolationS16": (None, [P, SS, P, SS, I32, U32, U32, U32]),
        "or_ChromaInterpolation": (None, [P, SS, P, SS, I32, U32, U32]),
        "or_ChromaInterpolationS16": (None, [P, SS, P, SS, I32, U32, U32]),
        "or_RefineFracBlocks": (None, [P, P, P, SS, SS, P, P, P, SS, SS, I32, I32, P, I32,
                                       ctypes.POINTER(OrRefinePrm), P, P, P, P, P, P, P, P, SS, SS]),
        "or_SearchBlocks": (None, [P, SS, P, SS, I32, I32, P, I32, I32, I32, U32, I32]),
        "or_FullSearchFrame": (None, [P, SS, P, SS, I32, I32, I32, I32, I32, U32, I32, I32, I32,
                                      P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib
