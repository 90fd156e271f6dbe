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


def _text(p, n):
    """The n bytes of UTF-8 text at p (a C pointer or address) as a str, decoded in place."""
    if n == 0:
        return ""
    addr = ctypes.cast(p, ctypes.c_void_p).value
    return str(memoryview((ctypes.c_char * n).from_address(addr)).cast("B"), "utf-8")


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
        text = _text(buf, n.value)
        self.lib.rv_free(buf)
        return text

    def fisher(self, tables):
        """Device two-sided Fisher test on an (n, 4) integer array of 2x2 tables (n11 n12 n21 n22);
        returns an (n, 3) float64 array of (two-sided p, left tail, right tail)."""
        t = np.ascontiguousarray(tables, np.int32).reshape(-1, 4)
        out = np.zeros((len(t), 3), np.float64)
        st = self.lib.rv_debug_fisher(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(t),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if st != 0:
            raise self._error(st)
        return out

    def slot_order(self, keys, transient):
        """The device's robin_hood slot order after inserting the distinct `keys` in order into an
        empty map, erasing each key whose `transient` flag is set right after its insertion.
        Returns the keys present at the end in slot order, or None when the device path declines
        the sequence."""
        k = np.ascontiguousarray(keys, np.int32)
        t = np.ascontiguousarray(transient, np.uint8)
        if len(k) != len(t):
            raise ValueError("keys and transient flags differ in length")
        out = np.zeros(max(1, len(k)), np.int32)
        n_out = ctypes.c_size_t()
        on = ctypes.c_int()
        st = self.lib.rv_debug_slot_order(self.h, k.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(k),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(n_out),
                                          ctypes.byref(on))
        if st != 0:
            raise self._error(st)
        return out[:n_out.value].copy() if on.value else None

    def find35(self, pairs):
        """The device's find35match on a list of (seq5, seq3) string pairs -> list of (b5, b3, score)."""
        arena, idx = bytearray(), []
        for a, b in pairs:
            idx += [len(arena), len(a)]
            arena += a.encode()
            idx += [len(arena), len(b)]
            arena += b.encode()
        p = np.ascontiguousarray(idx, np.int32)
        out = np.zeros(3 * len(pairs), np.int32)
        st = self.lib.rv_debug_find35(self.h, bytes(arena), len(arena), p.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      len(pairs), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if st != 0:
            raise self._error(st)
        return [tuple(int(x) for x in out[3 * k:3 * k + 3]) for k in range(len(pairs))]

    def run_region(self, chrom, start, end, reads, contig_name, contig_bases, gene=None):
        region, r, contig = self._args(chrom, start, end, reads, contig_name, contig_bases, gene)
        cs = ctypes.c_void_p()
        st = self.lib.rv_run_region(self.h, ctypes.byref(region), ctypes.byref(r), ctypes.byref(contig),
                                    ctypes.byref(cs))
        if st != 0:
            raise self._error(st)
        n = ctypes.c_size_t()
        p = self.lib.rv_callset_text(cs, ctypes.byref(n))
        text = _text(p, n.value)
        self.lib.rv_callset_free(cs)
        return text

    def run_region_calls(self, chrom, start, end, reads, contig_name, contig_bases, gene=None):
        """run_region, returning (text, calls): the structured records (rv_call) as dicts, in the
        text's line order."""
        region, r, contig = self._args(chrom, start, end, reads, contig_name, contig_bases, gene)
        cs = ctypes.c_void_p()
        st = self.lib.rv_run_region(self.h, ctypes.byref(region), ctypes.byref(r), ctypes.byref(contig),
                                    ctypes.byref(cs))
        if st != 0:
            raise self._error(st)
        try:
            n = ctypes.c_size_t()
            text = _text(self.lib.rv_callset_text(cs, ctypes.byref(n)), n.value)
            calls = []
            for i in range(self.lib.rv_callset_count(cs)):
                c = ctypes.cast(self.lib.rv_callset_get(cs, i), ctypes.POINTER(_lib.RvCall)).contents
                d = {k: getattr(c, k) for k, _ in c._fields_}
                for k in ("refallele", "varallele", "genotype", "bias", "leftseq", "rightseq", "vartype"):
                    d[k] = d[k].decode() if d[k] is not None else None
                calls.append(d)
        finally:
            self.lib.rv_callset_free(cs)
        return text, calls

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
        text = _text(p, n.value)
        self.lib.rv_callset_free(cs)
        return text

    def set_screen(self, params, threshold):
        """Screened mode: the realigner skips candidates the network scores below `threshold`
        (threshold <= 0 or params None: exact mode)."""
        if params is None:
            st = self.lib.rv_set_screen(self.h, None, 0, 0.0)
        else:
            self._screen = np.ascontiguousarray(params, np.float32)
            st = self.lib.rv_set_screen(self.h, self._screen.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        self._screen.size, float(threshold))
        if st != 0:
            raise self._error(st)

    def screen_dataset(self, chrom, start, end, reads, contig_name, contig_bases):
        """Runs the region exactly and returns its realignment candidates' features [n, 32], change
        labels [n] (1: the realigner changed a count there) and kinds [n]."""
        region, r, contig = self._args(chrom, start, end, reads, contig_name, contig_bases)
        fp = ctypes.POINTER(ctypes.c_float)()
        lp = ctypes.POINTER(ctypes.c_uint8)()
        kp = ctypes.POINTER(ctypes.c_int32)()
        n = ctypes.c_size_t()
        st = self.lib.rv_screen_dataset(self.h, ctypes.byref(region), ctypes.byref(r), ctypes.byref(contig),
                                        ctypes.byref(fp), ctypes.byref(lp), ctypes.byref(kp), ctypes.byref(n))
        if st != 0:
            raise self._error(st)
        m = n.value
        if m == 0:
            return np.zeros((0, 32), np.float32), np.zeros(0, np.uint8), np.zeros(0, np.int32)
        feats = np.ctypeslib.as_array(fp, shape=(m * 32,)).reshape(m, 32).copy()
        labels = np.ctypeslib.as_array(lp, shape=(m,)).copy()
        kinds = np.ctypeslib.as_array(kp, shape=(m,)).copy()
        for p in (fp, lp, kp):
            self.lib.rv_free(ctypes.cast(p, ctypes.c_void_p))
        return feats, labels, kinds

    def stats(self):
        s = _lib.RvStats()
        self.lib.rv_get_stats(self.h, ctypes.byref(s))
        out = {k: getattr(s, k) for k, _ in s._fields_}
        out["detail_ms"] = list(s.detail_ms)
        out["timeline_ms"] = list(s.timeline_ms)
        return out


class Mlp:
    """The realignment screen's network on the GPU (rv_mlp_*): training with the library's own
    forward/backward kernels, gradients and scores."""

    def __init__(self, device=0, params=None, seed=1):
        self.lib = _lib.load()
        self.n_params = self.lib.rv_mlp_param_count()
        h = ctypes.c_void_p()
        init = None
        if params is not None:
            self._init = np.ascontiguousarray(params, np.float32)
            assert self._init.size == self.n_params
            init = self._init.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        st = self.lib.rv_mlp_open(device, init, seed, ctypes.byref(h))
        if st != 0:
            raise EngineError(f"rv_mlp_open failed with status {st}")
        self.h = h

    def _check(self, st):
        if st != 0:
            msg = self.lib.rv_mlp_last_error(self.h)
            raise EngineError(f"status {st}: {msg.decode() if msg else ''}")

    def load(self, X, y=None):
        self._X = np.ascontiguousarray(X, np.float32)
        self._y = None if y is None else np.ascontiguousarray(y, np.uint8)
        fp = self._X.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        yp = None if self._y is None else self._y.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        self._check(self.lib.rv_mlp_load(self.h, fp, yp, self._X.shape[0]))

    def train(self, epochs, batch=1024, lr=1e-3, pos_weight=1.0, seed=1):
        losses = np.zeros(epochs, np.float32)
        self._check(self.lib.rv_mlp_train(self.h, epochs, batch, lr, pos_weight, seed,
                                          losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return losses

    def gradient(self, pos_weight=1.0):
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.rv_mlp_gradient(self.h, pos_weight, g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             ctypes.byref(loss)))
        return g, loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0], np.float32)
        self._check(self.lib.rv_mlp_predict(self.h, X.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), X.shape[0],
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._check(self.lib.rv_mlp_get_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p

    def close(self):
        if getattr(self, "h", None):
            self.lib.rv_mlp_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


class EnginePool:
    """Several engines on one device, one host thread each (rv_pool_*): a list of regions runs with
    one region's host stages overlapping others' device work; the texts come back in region order."""

    def __init__(self, workers=4, device=0, sample="S1", **cfg_fields):
        self.lib = _lib.load()
        cfg = _lib.RvConfig()
        self.lib.rv_config_default(ctypes.byref(cfg))
        for k, v in cfg_fields.items():
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
        st = self.lib.rv_pool_open(ctypes.byref(cfg), device, workers, ctypes.byref(h))
        if st != 0:
            raise EngineError(f"rv_pool_open failed with status {st}")
        self.h = h

    def set_screen(self, params, threshold):
        if params is None:
            st = self.lib.rv_pool_set_screen(self.h, None, 0, 0.0)
        else:
            self._screen = np.ascontiguousarray(params, np.float32)
            st = self.lib.rv_pool_set_screen(self.h, self._screen.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             self._screen.size, float(threshold))
        if st != 0:
            raise EngineError(f"rv_pool_set_screen failed with status {st}")

    def run_regions(self, regions, bam, contigs, bam2=None):
        """regions: [(chrom, start, end, gene)]; bam, bam2: rabbitvar_amd.BamFile (bam2 for somatic
        mode); contigs: {name: sequence bytes}.  Returns the texts, one per region."""
        n = len(regions)
        keep = []
        regs = (_lib.RvRegion * max(n, 1))()
        for i, (chrom, start, end, gene) in enumerate(regions):
            c, g = chrom.encode(), (gene or chrom).encode()
            keep += [c, g]
            regs[i] = _lib.RvRegion(c, start, end, g)
        names = list(contigs)
        cts = (_lib.RvContig * max(len(names), 1))()
        for i, name in enumerate(names):
            nb = name.encode()
            keep += [nb, contigs[name]]
            cts[i] = _lib.RvContig(nb, len(contigs[name]), contigs[name])
        out = (ctypes.c_void_p * max(n, 1))()
        st = self.lib.rv_pool_run(self.h, regs, n, bam.h, bam2.h if bam2 is not None else None, cts, len(names), out)
        if st != 0:
            msg = self.lib.rv_pool_last_error(self.h)
            raise EngineError(f"status {st}: {msg.decode() if msg else ''}")
        texts = []
        for i in range(n):
            ln = ctypes.c_size_t()
            p = self.lib.rv_callset_text(out[i], ctypes.byref(ln))
            texts.append(_text(p, ln.value))
            self.lib.rv_callset_free(out[i])
        return texts

    def close(self):
        if getattr(self, "h", None):
            self.lib.rv_pool_close(self.h)
            self.h = None

    def __del__(self):
        self.close()
