"""BAM decoding (rv_bam_*) against two things it does not share code with: Python's gzip module and
a struct-level parser written from the SAM/BAM specification here, and the same records' RVB file
written by rvsynth.  Region queries follow htslib's iterator: records of the contig that start before
the region's end and end (bam_endpos) after its start, in file order."""
import gzip
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import SYNTH

FIELDS = ["pos0", "flag", "mapq", "has_nm", "nm", "mpos0", "tid", "mtid", "seq_off", "cigar_off", "bases", "quals",
          "cigar"]


@pytest.fixture(scope="module")
def bamdir(tmp_path_factory, tools):
    d = str(tmp_path_factory.mktemp("bam"))
    subprocess.run([SYNTH, "--out", d, "--seed", "3", "--len", "60000", "--depth", "30", "--bam", "--dup-rate", "0.1",
                    "--single-end-frac", "0.2", "--unmapped-frac", "0.05", "--indel-rate", "0.001"], check=True,
                   capture_output=True)
    return d


def parse_bam(path):
    """Records as dicts, from gzip + the specification's record layout."""
    data = gzip.open(path, "rb").read()
    assert data[:4] == b"BAM\1"
    l_text = struct.unpack_from("<i", data, 4)[0]
    p = 8 + l_text
    n_ref = struct.unpack_from("<i", data, p)[0]
    p += 4
    refs = []
    for _ in range(n_ref):
        ln = struct.unpack_from("<i", data, p)[0]
        name = data[p + 4:p + 4 + ln - 1].decode()
        refs.append((name, struct.unpack_from("<i", data, p + 4 + ln)[0]))
        p += 8 + ln
    recs = []
    while p < len(data):
        bs, tid, pos, lrn, mapq, _bin, ncig, flag, lseq, mtid, mpos, _tlen = struct.unpack_from("<iiiBBHHHiiii", data, p)
        q = p + 36 + lrn
        cig = list(struct.unpack_from(f"<{ncig}I", data, q))
        q += 4 * ncig
        packed = data[q:q + (lseq + 1) // 2]
        seq = "".join("=ACMGRSVTWYHKDBN"[(packed[i // 2] >> (4 if i % 2 == 0 else 0)) & 15] for i in range(lseq))
        q += (lseq + 1) // 2
        qual = list(data[q:q + lseq])
        q += lseq
        nm = None
        end = p + 4 + bs
        while q < end:
            tag, ty = data[q:q + 2], chr(data[q + 2])
            q += 3
            if ty == "i":
                v = struct.unpack_from("<i", data, q)[0]
                q += 4
            else:
                raise AssertionError(ty)
            if tag == b"NM":
                nm = v
        recs.append(dict(tid=tid, pos0=pos, mapq=mapq, flag=flag, mtid=mtid, mpos0=mpos, cigar=cig, seq=seq,
                         qual=qual, nm=nm))
        p = end
    return refs, recs


# None: the library's default (this small file is inflated whole); 0: the streamed mode (block
# index, each query inflates only its blocks)
MODES = [None, 0]


@pytest.mark.parametrize("cache", MODES)
def test_decode_matches_spec_parser_and_rvb(bamdir, cache):
    from rabbitvar_amd import BamFile, read_rvb

    refs, recs = parse_bam(os.path.join(bamdir, "reads.bam"))
    b = BamFile(os.path.join(bamdir, "reads.bam"), threads=4, max_cached_bytes=cache)
    assert b.targets() == refs == [("chrS", 60000)]
    assert b.n_records() == len(recs) > 10000
    got = b.query("chrS", 1, 60000)
    assert got.n == len(recs)
    for i in range(0, len(recs), 97):
        r = recs[i]
        assert got.pos0[i] == r["pos0"] and got.flag[i] == r["flag"] and got.mapq[i] == r["mapq"]
        assert got.mpos0[i] == r["mpos0"] and got.mtid[i] == r["mtid"] and got.tid[i] == r["tid"]
        assert got.has_nm[i] == 1 and got.nm[i] == r["nm"]
        s0, s1 = got.seq_off[i], got.seq_off[i + 1]
        assert bytes(got.bases[s0:s1]).decode() == r["seq"] and list(got.quals[s0:s1]) == r["qual"]
        c0, c1 = got.cigar_off[i], got.cigar_off[i + 1]
        assert list(got.cigar[c0:c1]) == r["cigar"]
    want = read_rvb(os.path.join(bamdir, "reads.rvb"))
    for k in FIELDS:
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    b.close()


@pytest.mark.parametrize("cache", MODES)
@pytest.mark.parametrize("start,end", [(1, 1), (1, 150), (20001, 20001), (30000, 30500), (59850, 60000),
                                       (60001, 70000)])
def test_region_queries(bamdir, start, end, cache):
    from rabbitvar_amd import BamFile, read_rvb

    b = BamFile(os.path.join(bamdir, "reads.bam"), threads=3, max_cached_bytes=cache)
    got = b.query("chrS", start, end)
    want = read_rvb(os.path.join(bamdir, "reads.rvb")).select(start, end)
    for k in FIELDS:
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    # the rule itself, from the spec parser's records
    _, recs = parse_bam(os.path.join(bamdir, "reads.bam"))
    exp = []
    for r in recs:
        rl = 0 if r["flag"] & 4 else sum(c >> 4 for c in r["cigar"] if (c & 15) in (0, 2, 3, 7, 8))
        endpos = r["pos0"] + (rl if rl > 0 else 1)
        if r["pos0"] < end and endpos > start - 1:
            exp.append(r["pos0"])
    assert list(got.pos0) == exp


def test_threads_agree_and_errors(bamdir, tmp_path):
    from rabbitvar_amd import BamFile

    a = BamFile(os.path.join(bamdir, "reads.bam"), threads=1).query("chrS", 1, 60000)
    b = BamFile(os.path.join(bamdir, "reads.bam"), threads=8).query("chrS", 1, 60000)
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    with pytest.raises(ValueError):
        BamFile(os.path.join(bamdir, "reads.bam")).query("chrX", 1, 100)
    raw = open(os.path.join(bamdir, "reads.bam"), "rb").read()
    bad = str(tmp_path / "truncated.bam")
    open(bad, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(OSError):
        BamFile(bad)
    flip = bytearray(raw)
    flip[5000] ^= 0xFF  # inside a compressed block: the CRC check or inflate fails
    bad2 = str(tmp_path / "corrupt.bam")
    open(bad2, "wb").write(bytes(flip))
    with pytest.raises(OSError):
        BamFile(bad2)
    for path in (bad, bad2):  # the streamed mode finds the same damage in its indexing pass
        with pytest.raises(OSError):
            BamFile(path, max_cached_bytes=0)
    with pytest.raises(OSError):
        BamFile(str(tmp_path / "missing.bam"))


def write_bgzf(path, data, block=60000):
    """BGZF from the specification: gzip members of at most 64 KiB input with a "BC" extra field
    holding the member size - 1, then the empty end-of-file member."""
    import zlib

    out = bytearray()

    def member(chunk):
        c = zlib.compressobj(6, zlib.DEFLATED, -15)
        body = c.compress(chunk) + c.flush()
        bsize = 12 + 6 + len(body) + 8
        out.extend(b"\x1f\x8b\x08\x04" + b"\0" * 4 + b"\0\xff" + struct.pack("<H", 6) + b"BC" +
                   struct.pack("<HH", 2, bsize - 1) + body + struct.pack("<II", zlib.crc32(chunk), len(chunk)))

    for i in range(0, len(data), block):
        member(data[i:i + block])
    member(b"")
    open(path, "wb").write(bytes(out))


def record_spans(data):
    l_text = struct.unpack_from("<i", data, 4)[0]
    p = 8 + l_text
    n_ref = struct.unpack_from("<i", data, p)[0]
    p += 4
    for _ in range(n_ref):
        p += 8 + struct.unpack_from("<i", data, p)[0]
    head, spans = data[:p], []
    while p < len(data):
        bs = struct.unpack_from("<i", data, p)[0]
        spans.append((p, p + 4 + bs))
        p += 4 + bs
    return head, spans


@pytest.mark.parametrize("blocksize", [60000, 977])
def test_unsorted_and_small_blocks_both_modes(bamdir, tmp_path, blocksize):
    """An unsorted file (records in reverse order) and records split over many small blocks: both
    modes return the same records, in file order, as the region rule applied to the spec parser's
    records."""
    from rabbitvar_amd import BamFile

    data = gzip.open(os.path.join(bamdir, "reads.bam"), "rb").read()
    head, spans = record_spans(data)
    rev = head + b"".join(data[a:b] for a, b in reversed(spans))
    paths = {"rev": str(tmp_path / "rev.bam"), "fwd": str(tmp_path / "fwd.bam")}
    write_bgzf(paths["rev"], rev, blocksize)
    write_bgzf(paths["fwd"], data, blocksize)
    for name, path in paths.items():
        _, recs = parse_bam(path)
        a = BamFile(path, threads=3)
        z = BamFile(path, threads=3, max_cached_bytes=0)
        assert a.n_records() == z.n_records() == len(recs)
        for start, end in [(1, 60000), (30000, 30500), (59850, 60000), (1, 1)]:
            ga, gz = a.query("chrS", start, end), z.query("chrS", start, end)
            for k in FIELDS:
                assert np.array_equal(getattr(ga, k), getattr(gz, k)), (name, start, end, k)
            exp = []
            for r in recs:
                rl = 0 if r["flag"] & 4 else sum(c >> 4 for c in r["cigar"] if (c & 15) in (0, 2, 3, 7, 8))
                if r["pos0"] < end and r["pos0"] + (rl if rl > 0 else 1) > start - 1:
                    exp.append(r["pos0"])
            assert list(gz.pos0) == exp, (name, start, end)
        a.close()
        z.close()
