"""Command line: the calling path with the reference's option names (src/Launcher.cpp:299-494 for
the options this path reads; region planning as RegionBuilder for -R and -i, src/RegionBuilder.cpp).

    python -m rabbitvar_amd -G ref.fa -b sample.bam -R chr1:1000-2000 -N S1 --fisher [--out calls.txt]
    python -m rabbitvar_amd -G ref.fa -b 'tumor.bam|normal.bam' -N 'T|N' -i panel.bed --fisher

Reads come from BAM (decoded by the library) or from RVB read-batch files.  The output is the text
the reference prints, region by region in the order given.  --screen turns on the learned
realignment screen (opt-in; the default is the exact path).
"""
import argparse
import sys

OPTIONS = [  # (flag, config field, type)
    ("-f", "freq", float), ("-q", "goodq", float), ("-o", "qratio", float), ("-O", "mapq", float),
    ("-V", "lofreq", float), ("-X", "vext", int), ("-m", "mismatch", int), ("-T", "trimBasesAfter", int),
    ("-k", "performLocalRealignment", int), ("-I", "indelsize", int), ("-B", "minBiasReads", int),
    ("-r", "minr", int), ("-P", "readPosFilter", int), ("-M", "minmatch", int), ("-L", "SVMINLEN", int),
    ("-Y", "referenceExtension", int), ("-x", "numberNucleotideToExtend", int), ("-Q", "mappingQuality", int),
    ("-J", "crisprCuttingSite", int),
]


def fasta_names(path):
    with open(path, "rb") as f:
        return {ln[1:].split()[0].decode() for ln in f if ln.startswith(b">")}


class Reads:
    def __init__(self, path, threads):
        from . import BamFile, read_rvb

        self.bam = None
        self.batch = None
        if path.endswith(".rvb"):
            self.batch = read_rvb(path)
        else:
            self.bam = BamFile(path, threads=threads)

    def region(self, chrom, start, end):
        if self.bam is not None:
            return self.bam.query(chrom, start, end)
        return self.batch.select(start, end)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m rabbitvar_amd", description=__doc__.splitlines()[0])
    ap.add_argument("-G", required=True, help="reference FASTA")
    ap.add_argument("-b", required=True, help="BAM (or .rvb); 'first|second' for somatic mode")
    ap.add_argument("-N", default="", help="sample name ('T|N' in somatic mode)")
    ap.add_argument("-R", help="region chr:start-end[:gene]")
    ap.add_argument("-i", "--bed", help="BED file of regions")
    ap.add_argument("-c", type=int, help="BED column of the chromosome (1-based)")
    ap.add_argument("-S", type=int, help="BED column of the region start")
    ap.add_argument("-E", type=int, help="BED column of the region end")
    ap.add_argument("-s", type=int, help="BED column of the segment starts")
    ap.add_argument("-e", type=int, help="BED column of the segment ends")
    ap.add_argument("-g", type=int, help="BED column of the gene name")
    ap.add_argument("-d", default="\t", help="BED delimiter")
    ap.add_argument("-z", action="store_true", help="zero-based coordinates")
    ap.add_argument("--auto_resize", action="store_true", help="split BED regions longer than 10 kb")
    ap.add_argument("-F", default="0x504", help="flag filter (hex)")
    ap.add_argument("-t", action="store_true", help="duplicate filter")
    ap.add_argument("-Z", type=float, dest="downsample",
                    help="downsampling fraction: each read passing -F is dropped when the C library's rand() "
                         "draw is <= RAND_MAX * fraction (draws in region order)")
    ap.add_argument("-p", action="store_true", help="pileup mode: print every covered position")
    ap.add_argument("-3", dest="move3", action="store_true")
    ap.add_argument("-u", action="store_true")
    ap.add_argument("--UN", action="store_true")
    ap.add_argument("--fisher", action="store_true")
    ap.add_argument("--out", help="output file (default: stdout)")
    ap.add_argument("--th", type=int, default=8, help="BAM inflate threads")
    ap.add_argument("--workers", type=int, default=4, help="engines working on regions at once (BAM input)")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--screen", action="store_true", help="screened mode with the trained model")
    for flag, field, typ in OPTIONS:
        ap.add_argument(flag, dest=field, type=typ)
    a = ap.parse_args(argv)

    from . import read_fasta
    from .engine import Engine

    cfg = {field: getattr(a, field) for _, field, _ in OPTIONS if getattr(a, field) is not None}
    cfg.update(samfilter=int(a.F, 16), removeDuplicatedReads=int(a.t), moveIndelsTo3=int(a.move3), doPileup=int(a.p),
               uniqueModeAlignmentEnabled=int(a.u), uniqueModeSecondInPairEnabled=int(a.UN), fisher=int(a.fisher))
    if a.downsample is not None:
        cfg.update(isDownsampling=1, downsampling=a.downsample)
    from . import regions as rg

    extend = cfg.get("numberNucleotideToExtend", 0)
    paths = a.b.split("|")
    samples = [Reads(p, a.th) for p in paths]
    names = fasta_names(a.G)
    if samples[0].bam is not None:
        names |= {n for n, _ in samples[0].bam.targets()}
    if a.R:
        segs = rg.from_option(a.R, names, extend, a.z)
    elif a.bed:
        with open(a.bed) as f:
            segs = rg.from_bed(f.readlines(), names, {"c": a.c, "S": a.S, "E": a.E, "s": a.s, "e": a.e, "g": a.g},
                               extend, a.z, a.d)
        if a.auto_resize:
            segs = rg.auto_resize(segs)
    else:
        ap.error("give -R or -i")
    # a region that ends before it starts reads no records and prints nothing
    regions = [r for r in rg.flatten(segs) if r[1] <= r[2]]
    contigs = {}
    for chrom, _, _, _ in regions:
        if chrom not in contigs:
            contigs[chrom] = read_fasta(a.G, chrom)[1]
    out = open(a.out, "w") if a.out else sys.stdout
    if all(s.bam is not None for s in samples):
        # BAM input: the native pool reads each region's records and runs the regions concurrently
        from .engine import EnginePool

        pool = EnginePool(workers=a.workers, device=a.device, sample=a.N, **cfg)
        try:
            if a.screen:
                from . import screen

                params, threshold, _ = screen.load_model()
                pool.set_screen(params, threshold)
            texts = pool.run_regions(regions, samples[0].bam, contigs,
                                     bam2=samples[1].bam if len(samples) == 2 else None)
            for t in texts:
                out.write(t)
        finally:
            if a.out:
                out.close()
            pool.close()
        return 0
    eng = Engine(device=a.device, sample=a.N, **cfg)
    if a.screen:
        from . import screen

        params, threshold, _ = screen.load_model()
        eng.set_screen(params, threshold)
    try:
        for chrom, start, end, gene in regions:
            seq = contigs[chrom]
            rs = [s.region(chrom, start, end) for s in samples]
            if len(rs) == 2:
                text = eng.run_region_somatic(chrom, start, end, rs[0], rs[1], chrom, seq, gene=gene)
            else:
                text = eng.run_region(chrom, start, end, rs[0], chrom, seq, gene=gene)
            out.write(text)
    finally:
        if a.out:
            out.close()
        eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
