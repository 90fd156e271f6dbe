"""ctypes bindings for librabbitvar_amd.so (declarations in include/rabbitvar_amd.h).

The library is loaded from rabbitvar_amd/_build/.  There is no fallback: when the native library is
missing the import of the engine raises, naming the build command.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librabbitvar_amd.so")

c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_char_p = ctypes.c_char_p
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t


class RvConfig(ctypes.Structure):
    _fields_ = [
        ("goodq", c_double), ("bias", c_double), ("freq", c_double), ("qratio", c_double),
        ("mapq", c_double), ("lofreq", c_double),
        ("vext", c_int), ("mismatch", c_int), ("trimBasesAfter", c_int),
        ("performLocalRealignment", c_int), ("indelsize", c_int), ("minBiasReads", c_int),
        ("minr", c_int), ("moveIndelsTo3", c_int), ("samfilter", c_int), ("readPosFilter", c_int),
        ("doPileup", c_int), ("minmatch", c_int), ("uniqueModeAlignmentEnabled", c_int),
        ("uniqueModeSecondInPairEnabled", c_int), ("fisher", c_int), ("SVMINLEN", c_int),
        ("referenceExtension", c_int), ("numberNucleotideToExtend", c_int),
        ("mappingQuality", c_int), ("removeDuplicatedReads", c_int), ("crisprCuttingSite", c_int),
        ("sample", c_char_p), ("isDownsampling", c_int), ("downsampling", c_double),
    ]


class RvReads(ctypes.Structure):
    _fields_ = [
        ("n_reads", c_int64), ("pos0", c_void_p), ("flag", c_void_p), ("mapq", c_void_p),
        ("has_nm", c_void_p), ("nm", c_void_p), ("mpos0", c_void_p), ("tid", c_void_p),
        ("mtid", c_void_p), ("seq_off", c_void_p), ("cigar_off", c_void_p), ("bases", c_void_p),
        ("quals", c_void_p), ("cigar", c_void_p), ("on_device", c_int),
    ]


class RvCall(ctypes.Structure):
    _fields_ = [
        ("position", c_int32), ("start", c_int32), ("end", c_int32),
        ("total_pos_coverage", c_int32), ("position_coverage", c_int32),
        ("ref_fwd", c_int32), ("ref_rev", c_int32), ("var_fwd", c_int32), ("var_rev", c_int32),
        ("shift3", c_int32), ("msint", c_int32), ("hicnt", c_int32), ("hicov", c_int32),
        ("frequency", c_double), ("mean_position", c_double), ("mean_quality", c_double), ("mean_mapq", c_double),
        ("hi_lo_ratio", c_double), ("hi_freq", c_double), ("extra_freq", c_double), ("msi", c_double), ("nm", c_double),
        ("fisher_p", c_double), ("fisher_ratio", c_double), ("pstd", c_int32), ("qstd", c_int32),
        ("refallele", c_char_p), ("varallele", c_char_p), ("genotype", c_char_p), ("bias", c_char_p),
        ("leftseq", c_char_p), ("rightseq", c_char_p), ("vartype", c_char_p),
    ]


class RvRegion(ctypes.Structure):
    _fields_ = [("chr", c_char_p), ("start", c_int32), ("end", c_int32), ("gene", c_char_p)]


class RvContig(ctypes.Structure):
    _fields_ = [("name", c_char_p), ("length", c_int32), ("bases", c_char_p)]


class RvStats(ctypes.Structure):
    _fields_ = [
        ("prep_ms", c_double), ("walk_ms", c_double), ("anchor_ms", c_double), ("score_ms", c_double),
        ("total_device_ms", c_double), ("host_ms", c_double),
        ("reads_in", c_int64), ("reads_walked", c_int64), ("events", c_int64), ("clips", c_int64),
        ("ref_key_creations", c_int64), ("walk_errors", c_int64), ("dense_rows", c_int64),
        ("event_bytes", c_int64), ("cx_positions", c_int64), ("rows_fetched", c_int64),
        ("reduce_ms", c_double), ("realign_ms", c_double), ("order_ms", c_double), ("format_ms", c_double),
        ("wall_ms", c_double), ("order_sorted", c_int64), ("detail_ms", c_double * 16),
        ("call_ms", c_double), ("screen_ms", c_double), ("screen_candidates", c_int64),
        ("screen_skipped", c_int64), ("screen_list_ms", c_double),
        ("realign_device_pairs", c_int64), ("realign_device_ms", c_double),
        ("order_ops_ms", c_double), ("order_device_ms", c_double), ("order_ops", c_int64),
        ("timeline_ms", c_double * 16),
    ]


# exported symbols and their signatures: name -> (restype, argtypes)
SIGNATURES = {
    "rv_config_default": (None, [ctypes.POINTER(RvConfig)]),
    "rv_open": (c_int, [ctypes.POINTER(RvConfig), c_int, ctypes.POINTER(c_void_p)]),
    "rv_run_region": (c_int, [c_void_p, ctypes.POINTER(RvRegion), ctypes.POINTER(RvReads),
                              ctypes.POINTER(RvContig), ctypes.POINTER(c_void_p)]),
    "rv_run_region_somatic": (c_int, [c_void_p, ctypes.POINTER(RvRegion), ctypes.POINTER(RvReads),
                                      ctypes.POINTER(RvReads), ctypes.POINTER(RvContig),
                                      ctypes.POINTER(c_void_p)]),
    "rv_callset_count": (c_size_t, [c_void_p]),
    "rv_callset_get": (c_void_p, [c_void_p, c_size_t]),
    "rv_callset_text": (ctypes.POINTER(ctypes.c_char), [c_void_p, ctypes.POINTER(c_size_t)]),
    "rv_callset_free": (None, [c_void_p]),
    "rv_get_stats": (c_int, [c_void_p, ctypes.POINTER(RvStats)]),
    "rv_last_error": (c_char_p, [c_void_p]),
    "rv_close": (None, [c_void_p]),
    "rv_debug_pileup": (c_int, [c_void_p, ctypes.POINTER(RvRegion), ctypes.POINTER(RvReads),
                                ctypes.POINTER(RvContig), ctypes.POINTER(ctypes.POINTER(ctypes.c_char)),
                                ctypes.POINTER(c_size_t)]),
    "rv_free": (None, [c_void_p]),
    "rv_debug_fisher": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_int32), c_size_t, ctypes.POINTER(ctypes.c_double)]),
    "rv_debug_find35": (c_int, [c_void_p, ctypes.c_char_p, c_size_t, ctypes.POINTER(ctypes.c_int32), c_size_t,
                                ctypes.POINTER(ctypes.c_int32)]),
    "rv_debug_slot_order": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint8), c_size_t,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(c_size_t), ctypes.POINTER(c_int)]),
    "rv_fisher_tables_noalloc": (c_int, [ctypes.POINTER(ctypes.c_int32), c_size_t, ctypes.POINTER(ctypes.c_double)]),
    "rv_set_screen": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), c_size_t, ctypes.c_float]),
    "rv_screen_dataset": (c_int, [c_void_p, ctypes.POINTER(RvRegion), ctypes.POINTER(RvReads),
                                  ctypes.POINTER(RvContig), ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                  ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                  ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)), ctypes.POINTER(c_size_t)]),
    "rv_mlp_param_count": (c_size_t, []),
    "rv_mlp_open": (c_int, [c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_uint64, ctypes.POINTER(c_void_p)]),
    "rv_mlp_load": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint8), c_size_t]),
    "rv_mlp_train": (c_int, [c_void_p, c_int, c_int, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                             ctypes.POINTER(ctypes.c_float)]),
    "rv_mlp_gradient": (c_int, [c_void_p, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float)]),
    "rv_mlp_predict": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), c_size_t, ctypes.POINTER(ctypes.c_float)]),
    "rv_mlp_get_params": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "rv_mlp_last_error": (c_char_p, [c_void_p]),
    "rv_mlp_close": (None, [c_void_p]),
    "rv_bam_open": (c_int, [c_char_p, c_int, ctypes.POINTER(c_void_p)]),
    "rv_bam_open_bounded": (c_int, [c_char_p, c_int, ctypes.c_uint64, ctypes.POINTER(c_void_p)]),
    "rv_bam_last_error": (c_char_p, [c_void_p]),
    "rv_bam_n_targets": (ctypes.c_int32, [c_void_p]),
    "rv_bam_target_name": (c_char_p, [c_void_p, ctypes.c_int32]),
    "rv_bam_target_len": (ctypes.c_int32, [c_void_p, ctypes.c_int32]),
    "rv_bam_n_records": (c_size_t, [c_void_p]),
    "rv_bam_query": (c_int, [c_void_p, c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(RvReads)]),
    "rv_bam_close": (None, [c_void_p]),
    "rv_pool_open": (c_int, [ctypes.POINTER(RvConfig), c_int, c_int, ctypes.POINTER(c_void_p)]),
    "rv_pool_run": (c_int, [c_void_p, ctypes.POINTER(RvRegion), c_size_t, c_void_p, c_void_p,
                            ctypes.POINTER(RvContig), c_size_t, ctypes.POINTER(c_void_p)]),
    "rv_pool_set_screen": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), c_size_t, ctypes.c_float]),
    "rv_pool_last_error": (c_char_p, [c_void_p]),
    "rv_pool_close": (None, [c_void_p]),
}

_lib = None


def load():
    """Loads the native library (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `python -m rabbitvar_amd.build`")
    # PyTorch (DeviceReads, the tests' device buffers) ships its own HIP runtime.  A process that
    # initialises the system runtime through this library first finds no GPU when PyTorch starts
    # later, so PyTorch is loaded first when it is installed and the library binds to its runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
