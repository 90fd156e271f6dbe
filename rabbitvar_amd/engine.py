"""Python host mirror of the engine's C ABI (include/rabbitvar_amd.h).

    eng = Engine(sample="S1", freq=0.01, performLocalRealignment=1)
    text = eng.run_region("chrS", 1, 200000, reads, contig_name, contig_bases)

Options use the reference's Configuration field names (include/Configuration.h:98-401).
"""
import ctypes

import numpy as np

from . import _lib


class EngineError(RuntimeError):
    pass


_ARRAYS = ["pos0", "flag", "mapq", "has_nm", "nm", "mpos0", "tid", "mtid", "seq_off", "cigar_off", "bases", "quals",
           "cigar"]


class DeviceReads:
    """A read batch resident in device memory (torch tensors on a HIP device hold the buffers; the
    engine reads them in place).  Built from a host ReadBatch."""

    def __init__(self, batch, device="cuda:0"):
        import torch

        self.n = batch.n
        self.t = {}
        for name in _ARRAYS:
            a = np.ascontiguousarray(getattr(batch, name))
            if a.dtype == np.uint16:
                a = a.view(np.int16)
            elif a.dtype == np.uint32:
                a = a.view(np.int32)
            if a.size == 0:
                a = np.zeros(1, a.dtype)
            self.t[name] = torch.from_numpy(a).to(device)
        torch.cuda.synchronize()

    def nbytes(self):
        return sum(int(t.numel() * t.element_size()) for t in self.t.values())

    def rv_reads(self):
        return _lib.RvReads(self.n, *[ctypes.c_void_p(self.t[k].data_ptr()) for k in _ARRAYS], 1)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else ctypes.c_void_p(0)


class Engine:
    def __init__(self, device=0, sample="", **options):
        self.lib = _lib.load()
        cfg = _lib.RvConfig()
        self.lib.rv_config_default(ctypes.byref(cfg))
        for k, v in options.items():
            if not hasattr(cfg, k):
                raise KeyError(f"unknown configuration field {k}")
            setattr(cfg, k, v)
        if cfg.doPileup:
            cfg.freq = -1.0
            cfg.minr = 0
        self._sample = sample.encode()
        cfg.sample = self._sample
        self.cfg = cfg
        h = ctypes.c_void_p()
        st = self.lib.rv_open(ctypes.byref(cfg), device, ctypes.byref(h))
        if st != 0:
            raise EngineError(f"rv_open failed with status {st} (no HIP device or runtime error)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.rv_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _reads(self, reads):
        if isinstance(reads, DeviceReads):
            return reads.rv_reads()
        arrays = [np.ascontiguousarray(getattr(reads, k)) for k in _ARRAYS]
        self._keep.append(arrays)
        return _lib.RvReads(reads.n, *[_ptr(a) for a in arrays], 0)

    def _args(self, chrom, start, end, reads, contig_name, contig_bases, gene=None):
        self._keep = []
        r = self._reads(reads)
        g = (gene if gene is not None else chrom).encode()
        c = chrom.encode()
        self._keep += [g, c, contig_bases]
        region = _lib.RvRegion(c, start, end, g)
        contig = _lib.RvContig(contig_name.encode(), len(contig_bases), contig_bases)
        return region, r, contig

    def _error(self, st):
        msg = self.lib.rv_last_error(self.h)
        return EngineError(f"status {st}: {msg.decode() if msg else ''}")

    def debug_pileup(self, chrom, start, end, reads, contig_name, contig_bases):
        region, r, contig = self._args(chrom, start, end, reads, contig_name, contig_bases)
        buf = ctypes.POINTER(ctypes.c_char)()
        n = ctypes.c_size_t()
        st = self.lib.rv_debug_pileup(self.h, ctypes.byref(region), ctypes.byref(r), ctypes.byref(contig),
                                      ctypes.byref(buf), ctypes.byref(n))
        if st != 0:
            raise self._error(st)
        text = ctypes.string_at(buf, n.value).decode()
        self.lib.rv_free(buf)
        return text

    def run_region(self, chrom, start, end, reads, contig_name, contig_bases, gene=None):
        region, r, contig = self._args(chrom, start, end, reads, contig_name, contig_bases, gene)
        cs = ctypes.c_void_p()
        st = self.lib.rv_run_region(self.h, ctypes.byref(region), ctypes.byref(r), ctypes.byref(contig),
                                    ctypes.byref(cs))
        if st != 0:
            raise self._error(st)
        n = ctypes.c_size_t()
        p = self.lib.rv_callset_text(cs, ctypes.byref(n))
        text = ctypes.string_at(p, n.value).decode()
        self.lib.rv_callset_free(cs)
        return text

    def run_region_somatic(self, chrom, start, end, tumor, normal, contig_name, contig_bases, gene=None):
        """Somatic mode for one region (the first and second samples of -b 'T|N'); returns the text
        the reference's somatic printer writes for it."""
        region, r1, contig = self._args(chrom, start, end, tumor, contig_name, contig_bases, gene)
        r2 = self._reads(normal)
        cs = ctypes.c_void_p()
        st = self.lib.rv_run_region_somatic(self.h, ctypes.byref(region), ctypes.byref(r1), ctypes.byref(r2),
                                            ctypes.byref(contig), ctypes.byref(cs))
        if st != 0:
            raise self._error(st)
        n = ctypes.c_size_t()
        p = self.lib.rv_callset_text(cs, ctypes.byref(n))
        text = ctypes.string_at(p, n.value).decode()
        self.lib.rv_callset_free(cs)
        return text

    def stats(self):
        s = _lib.RvStats()
        self.lib.rv_get_stats(self.h, ctypes.byref(s))
        out = {k: getattr(s, k) for k, _ in s._fields_}
        out["detail_ms"] = list(s.detail_ms)
        return out
