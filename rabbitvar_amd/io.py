"""Readers for the inputs of the calling path: single-contig FASTA and read-batch files (RVB1, the
structure-of-arrays layout written by rvsynth; see rabbitvar_amd/synth/rv_synth.cpp), plus the
region iterator's record selection."""
import numpy as np


def read_fasta(path, contig=None):
    """Returns (name, sequence bytes) of the first contig, or of the named one."""
    name = None
    parts = []
    with open(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                if name is not None and (contig is None or name == contig):
                    break
                name = line[1:].split()[0].decode()
                parts = []
                continue
            parts.append(line)
    if contig is not None and name != contig:
        raise KeyError(f"{path}: no contig {contig}")
    return name, b"".join(parts)


_FIELDS = [("pos0", np.int32, 0), ("flag", np.uint16, 0), ("mapq", np.uint8, 0), ("has_nm", np.uint8, 0),
           ("nm", np.int32, 0), ("mpos0", np.int32, 0), ("tid", np.int32, 0), ("mtid", np.int32, 0),
           ("seq_off", np.int64, 1), ("cigar_off", np.int64, 1)]


def _ranges(starts, lens):
    """Concatenation of arange(starts[i], starts[i] + lens[i]) for all i."""
    lens = np.asarray(lens, dtype=np.int64)
    total = int(lens.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    excl = np.concatenate([[0], np.cumsum(lens)[:-1]])
    return np.repeat(np.asarray(starts, np.int64) - excl, lens) + np.arange(total, dtype=np.int64)


class ReadBatch:
    """Aligned records as structure-of-arrays numpy arrays (sorted by pos0)."""

    def __init__(self, arrays):
        self.__dict__.update(arrays)
        self.n = int(len(self.pos0))

    def ref_len(self):
        """Reference span of each record (M/D/N/=/X lengths; 1 when empty or unmapped, as
        bam_endpos)."""
        ops = self.cigar & 0xF
        lens = (self.cigar >> 4).astype(np.int64)
        on_ref = np.isin(ops, (0, 2, 3, 7, 8))
        csum = np.concatenate([[0], np.cumsum(np.where(on_ref, lens, 0))])
        span = csum[self.cigar_off[1:]] - csum[self.cigar_off[:-1]]
        span = np.where(self.flag & 4, 0, span)
        return np.where(span > 0, span, 1)

    def select(self, start, end):
        """Records the region iterator returns for chr:start-end (1-based, inclusive): those whose
        [pos0, pos0 + span) overlaps [start - 1, end), in file order."""
        span = self.ref_len()
        keep = (self.pos0 < end) & (self.pos0 + span > start - 1)
        idx = np.nonzero(keep)[0]
        return self.take(idx)

    def take(self, idx):
        idx = np.asarray(idx, dtype=np.int64)
        so, co = self.seq_off, self.cigar_off
        slen = so[idx + 1] - so[idx]
        clen = co[idx + 1] - co[idx]
        new_so = np.concatenate([[0], np.cumsum(slen)]).astype(np.int64)
        new_co = np.concatenate([[0], np.cumsum(clen)]).astype(np.int64)
        bidx = _ranges(so[idx], slen)
        cidx = _ranges(co[idx], clen)
        arrays = {name: getattr(self, name)[idx] for name, _, extra in _FIELDS if not extra}
        arrays.update(seq_off=new_so, cigar_off=new_co, bases=self.bases[bidx], quals=self.quals[bidx],
                      cigar=self.cigar[cidx])
        return ReadBatch(arrays)


def write_rvb(path, batch):
    """Writes a ReadBatch in the RVB1 layout read_rvb() reads."""
    n = batch.n
    nb = int(batch.seq_off[-1]) if n else 0
    nc = int(batch.cigar_off[-1]) if n else 0
    with open(path, "wb") as f:
        f.write(b"RVB1")
        f.write(np.array([1], np.uint32).tobytes())
        f.write(np.array([n, nb, nc], np.int64).tobytes())
        arrays = [(getattr(batch, name), dtype) for name, dtype, _ in _FIELDS]
        arrays += [(batch.bases, np.uint8), (batch.quals, np.uint8), (batch.cigar, np.uint32)]
        for a, dtype in arrays:
            b = np.ascontiguousarray(a, dtype=dtype).tobytes()
            f.write(b)
            f.write(b"\0" * ((-len(b)) % 8))


def batch_from_records(records):
    """ReadBatch from a list of dicts with keys pos0, flag, mapq, nm (optional), mpos0, tid, mtid,
    seq (str), qual (list of Phred values) and cigar (list of (op, length), ops as BAM codes)."""
    n = len(records)
    arrays = {name: np.zeros(n + extra, dtype) for name, dtype, extra in _FIELDS}
    bases, quals, cigar = [], [], []
    for i, r in enumerate(records):
        arrays["pos0"][i] = r["pos0"]
        arrays["flag"][i] = r.get("flag", 0)
        arrays["mapq"][i] = r.get("mapq", 60)
        arrays["has_nm"][i] = 1 if "nm" in r else 0
        arrays["nm"][i] = r.get("nm", 0)
        arrays["mpos0"][i] = r.get("mpos0", -1)
        arrays["tid"][i] = r.get("tid", 0)
        arrays["mtid"][i] = r.get("mtid", -1)
        assert len(r["seq"]) == len(r["qual"])
        bases.extend(r["seq"].encode())
        quals.extend(r["qual"])
        cigar.extend((length << 4) | op for op, length in r["cigar"])
        arrays["seq_off"][i + 1] = len(bases)
        arrays["cigar_off"][i + 1] = len(cigar)
    arrays["bases"] = np.array(bases, np.uint8)
    arrays["quals"] = np.array(quals, np.uint8)
    arrays["cigar"] = np.array(cigar, np.uint32)
    return ReadBatch(arrays)


def read_rvb(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != b"RVB1":
        raise ValueError(f"{path}: not a read-batch file")
    n, nb, nc = np.frombuffer(data, dtype=np.int64, count=3, offset=8)
    off = 32
    arrays = {}

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(data, dtype=dtype, count=int(count), offset=off).copy()
        nbytes = a.nbytes
        off += nbytes + (-nbytes) % 8
        return a

    for name, dtype, extra in _FIELDS:
        arrays[name] = take(dtype, n + extra)
    arrays["bases"] = take(np.uint8, nb)
    arrays["quals"] = take(np.uint8, nb)
    arrays["cigar"] = take(np.uint32, nc)
    return ReadBatch(arrays)


class BamFile:
    """A BAM file decoded by the native library (rv_bam_*): query() returns the records of a region
    as a ReadBatch, as the reference's region iterator does.  Files whose inflated size is at most
    max_cached_bytes (default: the library's, 4 GiB) are inflated whole at open; larger ones keep a
    block index and inflate only the blocks a query needs."""

    def __init__(self, path, threads=8, max_cached_bytes=None):
        import ctypes

        from . import _lib

        self.lib = _lib.load()
        h = ctypes.c_void_p()
        if max_cached_bytes is None:
            st = self.lib.rv_bam_open(path.encode(), threads, ctypes.byref(h))
        else:
            st = self.lib.rv_bam_open_bounded(path.encode(), threads, int(max_cached_bytes), ctypes.byref(h))
        self.h = h
        if st != 0:
            msg = self.lib.rv_bam_last_error(h) if h.value else b""
            self.close()
            raise OSError(f"{path}: {msg.decode() if msg else 'cannot open'} (status {st})")

    def targets(self):
        return [(self.lib.rv_bam_target_name(self.h, t).decode(), self.lib.rv_bam_target_len(self.h, t))
                for t in range(self.lib.rv_bam_n_targets(self.h))]

    def n_records(self):
        return self.lib.rv_bam_n_records(self.h)

    def query(self, chrom, start, end):
        import ctypes

        from . import _lib

        r = _lib.RvReads()
        st = self.lib.rv_bam_query(self.h, chrom.encode(), start, end, ctypes.byref(r))
        if st != 0:
            msg = self.lib.rv_bam_last_error(self.h)
            raise ValueError(msg.decode() if msg else f"status {st}")
        n = r.n_reads

        def arr(ptr, dtype, count):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                         shape=(count,)).copy()

        so = arr(r.seq_off, np.int64, n + 1)
        co = arr(r.cigar_off, np.int64, n + 1)
        arrays = {"pos0": arr(r.pos0, np.int32, n), "flag": arr(r.flag, np.uint16, n),
                  "mapq": arr(r.mapq, np.uint8, n), "has_nm": arr(r.has_nm, np.uint8, n),
                  "nm": arr(r.nm, np.int32, n), "mpos0": arr(r.mpos0, np.int32, n), "tid": arr(r.tid, np.int32, n),
                  "mtid": arr(r.mtid, np.int32, n), "seq_off": so, "cigar_off": co,
                  "bases": arr(r.bases, np.uint8, int(so[-1]) if n else 0),
                  "quals": arr(r.quals, np.uint8, int(so[-1]) if n else 0),
                  "cigar": arr(r.cigar, np.uint32, int(co[-1]) if n else 0)}
        if n == 0:
            arrays["seq_off"] = np.zeros(1, np.int64)
            arrays["cigar_off"] = np.zeros(1, np.int64)
        return ReadBatch(arrays)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rv_bam_close(self.h)
        self.h = None

    def __del__(self):
        self.close()
