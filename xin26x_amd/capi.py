"""ctypes bindings of libxg266.so (include/xg266.h).

load() returns the library with every prototype set.  It raises if the HIP
library has not been built: there is no CPU fallback anywhere in the product.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libxg266.so")

P = ctypes.c_void_p
SS = ctypes.c_ssize_t
I16 = ctypes.c_int16
I32 = ctypes.c_int32
U32 = ctypes.c_uint32
U8 = ctypes.c_uint8

XG_MAX_CAND = 64
XG_LOOKAHEAD_STATIC_MAX_BYTES = 12 << 20      # include/xg266.h
XG_OK, XG_ERR_ARG, XG_ERR_HIP, XG_ERR_NOMEM, XG_ERR_BUDGET, XG_ERR_STATE = range(6)
MODE_PLAIN, MODE_BATCHED, MODE_GRAPH = 0, 1, 2


class XgPlane(ctypes.Structure):
    _fields_ = [("data", P), ("stride", SS), ("width", I32), ("height", I32), ("pad", I32)]


class XgPicture(ctypes.Structure):
    _fields_ = [("plane", XgPlane * 3)]


class XgBlock(ctypes.Structure):
    _fields_ = [("x", I16), ("y", I16), ("w", U8), ("h", U8), ("pic", U8), ("ref", U8),
                ("mvx", I32), ("mvy", I32), ("mvpx", I32), ("mvpy", I32)]


class XgRefineCfg(ctypes.Structure):
    _fields_ = [("ncand", I32), ("offx", I16 * XG_MAX_CAND), ("offy", I16 * XG_MAX_CAND),
                ("lambda_", U32), ("ctuSize", I32), ("useChroma", I32)]


class XgBlockResult(ctypes.Structure):
    _fields_ = [("mvx", I32), ("mvy", I32), ("cost", U32), ("cand", I32)]


class XgAmvpBlock(ctypes.Structure):
    _fields_ = [("x", I16), ("y", I16), ("w", U8), ("h", U8), ("pic", U8), ("ref", U8),
                ("ncand", I32), ("cand", I32 * 4)]


class XgAmvpResult(ctypes.Structure):
    _fields_ = [("mvpIdx", I32), ("sad", U32), ("mv", I32 * 2), ("meMv", I16 * 2)]


class XgServerCfg(ctypes.Structure):
    _fields_ = [("hbmBudget", ctypes.c_uint64), ("maxStreams", I32), ("maxBatchBlocks", I32),
                ("searchRange", I32), ("searchLambda", U32), ("maxDelayUs", I32), ("refine", XgRefineCfg),
                ("minBatchFrames", I32)]


XG_MEM_HOST, XG_MEM_HOST_PINNED, XG_MEM_DEVICE = range(3)


class XgServerFrame(ctypes.Structure):
    _fields_ = [("planesKind", I32), ("planes", ctypes.c_void_p * 3), ("strides", ctypes.c_ssize_t * 3),
                ("blocks", ctypes.c_void_p), ("nblocks", I32), ("outputsKind", I32), ("results", ctypes.c_void_p),
                ("pred", ctypes.c_void_p * 3), ("predStride", ctypes.c_ssize_t * 3), ("tag", ctypes.c_uint64)]


class XgPipelineCfg(ctypes.Structure):
    _fields_ = [("range", I32), ("sadLambda", U32), ("ctuSize", I32), ("refine", XgRefineCfg)]


class XgFrameResult(ctypes.Structure):
    _fields_ = [("stream", I32), ("status", I32), ("nblocks", I32), ("batchFrames", I32),
                ("tag", ctypes.c_uint64), ("submitNs", ctypes.c_int64), ("startNs", ctypes.c_int64),
                ("doneNs", ctypes.c_int64)]


class XgServerStats(ctypes.Structure):
    _fields_ = [("framesDone", ctypes.c_int64), ("framesDropped", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("bytesInUse", ctypes.c_uint64), ("bytesBudget", ctypes.c_uint64), ("streamsOpen", I32)]


class XgMv(ctypes.Structure):
    _fields_ = [("mvX", I16), ("mvY", I16)]


class XgDmvr(ctypes.Structure):
    _fields_ = [("ref0", P), ("stride0", SS), ("ref1", P), ("stride1", SS),
                ("cuPelX", I32), ("cuPelY", I32), ("width", U32), ("height", U32),
                ("mv0x", I32), ("mv0y", I32), ("mv1x", I32), ("mv1y", I32),
                ("minMvX", ctypes.c_int16), ("minMvY", ctypes.c_int16),
                ("maxMvX", ctypes.c_int16), ("maxMvY", ctypes.c_int16)]


class XgDmvrBlock(ctypes.Structure):
    _fields_ = [("x", I32), ("y", I32), ("width", I32), ("height", I32),
                ("mv0x", I32), ("mv0y", I32), ("mv1x", I32), ("mv1y", I32),
                ("minMvX", ctypes.c_int16), ("minMvY", ctypes.c_int16),
                ("maxMvX", ctypes.c_int16), ("maxMvY", ctypes.c_int16)]


class XgMeStruct(ctypes.Structure):
    _fields_ = [("width", U32), ("height", U32), ("minMv", XgMv), ("maxMv", XgMv),
                ("meCand", XgMv * 12), ("candNum", U32), ("bestMv", XgMv), ("predMv", XgMv),
                ("bestCost", U32), ("bestSad", U32), ("searchWidth", I32), ("searchHeight", I32),
                ("sadLambda", U32), ("input", P), ("inputStride", SS), ("ref", P), ("refStride", SS)]


class XgFuncStruct(ctypes.Structure):
    _fields_ = [("pfXinComputeSad", P * 8), ("pfXinComputeSatd", P * 8), ("pfXinComputeDist", P * 8),
                ("pfXinComputeAvgSad", P * 8), ("pfXinComputeAvgSatd", P * 8), ("pfXinComputeAvgDist", P * 8),
                ("pfXinComputeSadx8", P * 8), ("pfXinComputeAvgSadx8", P * 8),
                ("pfXinComputeSadx5", P * 8), ("pfXinComputeSadx3", P * 8),
                ("pfXinLumaInterp", P * 32), ("pfXinChromaInterp", P * 32),
                ("pfXinLumaInterpS16", P * 32), ("pfXinChromaInterpS16", P * 32),
                ("pfXinInterpAvg", P * 8), ("pfXinInterpWeight", P * 8),
                ("pfXinFor2dDct2", P * 64), ("pfXinFor1dDct", P * 24),
                ("pfXinBlockSubForDct", P * 8), ("pfXinQuantInvQuant", P * 8),
                ("pfXinInv2dSkip", P * 8), ("pfXinInv2dDct2", P * 64), ("pfXinInv1dDct", P * 24),
                ("pfXinBlockAddForDct", P * 8), ("pfXinBiliInterp", P * 32), ("pfXinComputeSadS16", P * 8),
                ("pfXinComputeHada8x8", P)]


# Every symbol include/xg266.h declares, with its prototype (restype, argtypes).
INTERP = (None, [P, SS, P, SS, I32, U32, U32, U32])
FDCT1D = (None, [P, SS, P, SS, I32, I32])
IDCT1D = (None, [P, SS, P, SS, I32, I32, I32])
BILI = (None, [P, SS, P, SS, I32, U32, U32])
PROTOS = {
    "xg_last_error": (I32, []),
    "xg_clear_error": (None, []),
    "xg_set_logger": (None, [P]),
    "xg_version": (ctypes.c_char_p, []),
    "Xg266ComputeSad": (None, [P, SS, P, SS, U32, U32, P]),
    "Xg266ComputeSatd": (None, [P, SS, P, SS, U32, U32, P]),
    "Xg266ComputeSadx8": (None, [P, SS, P, SS, U32, U32, P]),
    "Xg266ComputeSadx5": (None, [P, SS, P, P, P, P, P, SS, U32, U32, P]),
    "Xg266ComputeSadx3": (None, [P, SS, P, P, P, SS, U32, U32, P]),
    "Xg266ComputeHada8x8": (None, [P, SS, P]),
    "Xg266QuantInvQuant": (None, [P, P, P, SS, U32, U32, U32, U32, I32, I32, I32, I32, I32, I32, P]),
    "Xg266ComputeAvgSad": (None, [P, SS, P, P, SS, U32, U32, P]),
    "Xg266ComputeAvgSatd": (None, [P, SS, P, P, SS, U32, U32, P]),
    "Xg266ComputeAvgSadx8": (None, [P, SS, P, P, SS, U32, U32, P]),
    "Xg266InterpCopy": INTERP, "Xg266LumaInterpHor": INTERP, "Xg266LumaInterpVet": INTERP,
    "Xg266LumaInterpHorVet": INTERP, "Xg266InterpCopyU8S16": INTERP,
    "Xg266LumaInterpHorU8S16": INTERP, "Xg266LumaInterpVetU8S16": INTERP,
    "Xg266LumaInterpHorVetU8S16": INTERP, "Xg266LumaInterpVetS16U8": INTERP,
    "Xg266LumaInterpVetS16S16": INTERP, "Xg266ChromaInterpHor": INTERP,
    "Xg266ChromaInterpVet": INTERP, "Xg266ChromaInterpHorVet": INTERP,
    "Xg266ChromaInterpHorU8S16": INTERP, "Xg266ChromaInterpVetU8S16": INTERP,
    "Xg266ChromaInterpHorVetU8S16": INTERP, "Xg266ChromaInterpVetS16U8": INTERP,
    "Xg266ChromaInterpVetS16S16": INTERP,
    "Xg266InterpAvgS16U8": (None, [P, SS, P, SS, P, SS, U32, U32]),
    "Xg266InterpWeightS16U8": (None, [P, SS, P, SS, P, SS, U32, U32, I32, I32]),
    "Xg26xFDct2P2": FDCT1D, "Xg26xFDct2P4": FDCT1D, "Xg26xFDct2P8": FDCT1D,
    "Xg26xFDct2P16": FDCT1D, "Xg26xFDct2P32": FDCT1D, "Xg26xFDct2P64": FDCT1D,
    "Xg26xFDst7P4": FDCT1D, "Xg26xFDst7P8": FDCT1D, "Xg26xFDst7P16": FDCT1D, "Xg26xFDst7P32": FDCT1D,
    "Xg26xFDct8P4": FDCT1D, "Xg26xFDct8P8": FDCT1D, "Xg26xFDct8P16": FDCT1D, "Xg26xFDct8P32": FDCT1D,
    "Xg26xFDct2WxH": (None, [P, SS, P, SS, U32, U32, P, I32, I32]),
    "Xg26xBlockSubForDct": (None, [P, SS, P, SS, P, SS, U32, U32]),
    "Xg266FDctWxH": (I32, [P, SS, P, SS, U32, U32, U32, I32, I32]),
    "Xg26xIDct2P2": IDCT1D, "Xg26xIDct2P4": IDCT1D, "Xg26xIDct2P8": IDCT1D,
    "Xg26xIDct2P16": IDCT1D, "Xg26xIDct2P32": IDCT1D, "Xg26xIDct2P64": IDCT1D,
    "Xg26xIDst7P4": IDCT1D, "Xg26xIDst7P8": IDCT1D, "Xg26xIDst7P16": IDCT1D, "Xg26xIDst7P32": IDCT1D,
    "Xg26xIDct8P4": IDCT1D, "Xg26xIDct8P8": IDCT1D, "Xg26xIDct8P16": IDCT1D, "Xg26xIDct8P32": IDCT1D,
    "Xg266IDct2WxH": (None, [P, SS, P, SS, U32, U32, P, I32, I32]),
    "Xg266ISkipWxH": (None, [P, SS, P, SS, U32, U32]),
    "Xg266IDctWxH": (I32, [P, SS, P, SS, U32, U32, U32, I32, I32]),
    "Xg26xBlockRecon": (None, [P, SS, P, SS, P, SS, U32, U32]),
    "Xg266IDctAdd": (I32, [P, SS, P, SS, P, SS, U32, U32, U32, I32, I32]),
    "Xg266FuncInit": (I32, [ctypes.POINTER(XgFuncStruct), I32]),
    "Xg26xMotionEstFull": (None, [ctypes.POINTER(XgMeStruct)]),
    "Xg26xMotionEstDia": (None, [ctypes.POINTER(XgMeStruct)]),
    "Xg266FullSubPelMe": (None, [ctypes.POINTER(XgMeStruct)]),
    "xg_session_create": (I32, [ctypes.POINTER(P), I32]),
    "xg_session_destroy": (I32, [P]),
    "xg_session_stream": (P, [P]),
    "xg_session_set_picture": (I32, [P, I32, ctypes.POINTER(XgPicture)]),
    "xg_session_set_blocks": (I32, [P, P, I32]),
    "xg_session_set_tile": (I32, [P, I32, I32, I32]),
    "xg_session_get_tile": (I32, [P, I32, I32]),
    "xg_session_set_sched": (I32, [P, I32, I32]),
    "xg_session_autotune": (I32, [P, ctypes.POINTER(XgRefineCfg), I32, ctypes.c_char_p]),
    "xg_session_load_tiles": (I32, [P, ctypes.c_char_p]),
    "xg_refine_frac": (I32, [P, ctypes.POINTER(XgRefineCfg), ctypes.POINTER(XgPicture), I32]),
    "xg_session_outputs": (I32, [P, ctypes.POINTER(P), ctypes.POINTER(P)]),
    "xg_session_sync": (I32, [P]),
    "xg_session_read": (I32, [P, P, P]),
    "xg_session_kernel_time": (I32, [P, I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(I32)]),
    "xg_session_reset_timers": (I32, [P]),
    "xg_session_set_timing": (I32, [P, I32]),
    "xg_full_search_frame": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), I32, I32, I32,
                                   U32, I32, P, P, P]),
    "xg_session_set_pred": (I32, [P, I32, ctypes.POINTER(XgPicture)]),
    "xg_search_blocks": (I32, [P, I32, U32, I32]),
    "xg_server_create": (I32, [ctypes.POINTER(P), ctypes.POINTER(XgServerCfg)]),
    "xg_server_destroy": (I32, [P]),
    "xg_server_open_stream": (I32, [P, I32, I32, ctypes.POINTER(I32)]),
    "xg_server_close_stream": (I32, [P, I32]),
    "xg_server_submit": (I32, [P, I32, P, P, P, I32, P, P, P, ctypes.c_uint64]),
    "xg_server_submit_frame": (I32, [P, I32, ctypes.POINTER(XgServerFrame)]),
    "xg_server_poll": (I32, [P, ctypes.POINTER(XgFrameResult), I32, I32]),
    "xg_server_stats_get": (I32, [P, ctypes.POINTER(XgServerStats)]),
    "xg_me_pipeline": (I32, [P, ctypes.POINTER(XgPipelineCfg), ctypes.POINTER(XgPicture), ctypes.POINTER(XgPicture),
                             ctypes.POINTER(XgPicture), ctypes.POINTER(XgPicture), I32]),
    "xg_me_pipeline_read": (I32, [P, P, P, P, P, P, P]),
    "xg_fdct_plan": (I32, [P, P, P, I32]),
    "xg_fdct_run": (I32, [P, P, P]),
    "xg_recon_plan": (I32, [P, P, P, I32]),
    "xg_recon_run": (I32, [P, P, P, P]),
    "xg_quant_plan": (I32, [P, P, P, I32]),
    "xg_quant_run": (I32, [P, P, P, P, P]),
    "xg_fdct_quant_plan": (I32, [P, P, P, P, I32]),
    "xg_fdct_quant_run": (I32, [P, P, P, P, P]),
    "xg_lookahead_me": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), I32, I32, U32, P, P, P]),
    "xg_amvp_select": (I32, [P, P, I32, I32, I32, P]),
    "xg_mctf_subsample": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane)]),
    "xg_mctf_me_level": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), I32, I32, P, I32, I32, I32, I32,
                               P, P]),
    "xg_mctf_me": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), P, P]),
    "xg_amvp_from_field": (I32, [P, P, I32, I32, I32, P, I32, P]),
    "xg_lookahead_me_static": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), I32, I32, U32, P, P, P]),
    "xg_bipred_plan": (I32, [P, P, I32]),
    "xg_bipred_run": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), I32, P]),
    "xg_dmvr_plan": (I32, [P, P, I32, P]),
    "xg_dmvr_run": (I32, [P, ctypes.POINTER(XgPlane), ctypes.POINTER(XgPlane), P]),
    "Xg266BiliInterpCopy": BILI, "Xg266BiliInterpHor": BILI, "Xg266BiliInterpVet": BILI,
    "Xg266BiliInterpHorVet": BILI,
    "Xg26xComputeSadS16": (None, [P, SS, P, SS, U32, U32, U32, P]),
    "Xg266DeriveMvFromDmvr": (I32, [ctypes.POINTER(XgDmvr), P]),
}

_lib = None


def load(path=None):
    """Load libxg266.so and set every prototype.  Raises when it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    # torch wheels bundle their own HIP runtime (same soname as /opt/rocm's).
    # Whichever is loaded first serves the whole process; if libxg266.so were
    # loaded first, torch would later bind to a runtime it was not built
    # against and report no device.  Load torch's first when it is installed,
    # so the device buffers torch allocates and this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m xin26x_amd.build` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc, what=""):
    if rc != XG_OK:
        raise RuntimeError(f"{what} failed with xg status {rc}")
    return rc
