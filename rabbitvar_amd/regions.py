"""Region planning as the reference's launcher does it (src/RegionBuilder.cpp, src/Launcher.cpp):
-R chr:start-end[:gene], BED rows with -c/-S/-E/-s/-e/-g columns and comma-separated segments,
-x widening, -z zero-based starts, --auto_resize, and the "chr" prefix correction against the
contig names."""

# DEFAULT_BED_ROW_FORMAT (Launcher.cpp:21): 0-based columns of a refGene-style row
DEFAULT_COLUMNS = {"c": 2, "S": 6, "E": 7, "s": 9, "e": 10, "g": 12}
REGION_SIZE_MAX = 10000  # RegionBuilder.cpp:7


def correct_chromosome(chrom, names):
    """RegionBuilder::correctChromosome (:208-217): a name the contigs lack gains or loses "chr"."""
    if chrom in names:
        return chrom
    return chrom[3:] if chrom.startswith("chr") else "chr" + chrom


def columns(given):
    """Launcher.cpp:415-430: a given column is 1-based; an absent one is the default (0-based); -s/-e
    follow -S/-E when only those are given."""
    col = {k: (given[k] - 1 if given.get(k) is not None else DEFAULT_COLUMNS[k]) for k in DEFAULT_COLUMNS}
    if given.get("S") is not None and given.get("s") is None:
        col["s"] = col["S"]
    if given.get("E") is not None and given.get("e") is None:
        col["e"] = col["E"]
    return col


def from_option(text, names, extend=0, zero_based=False):
    """RegionBuilder::buildRegionFromConfiguration (:179-200)."""
    parts = text.split(":")
    chrom = correct_chromosome(parts[0], names)
    gene = parts[2] if len(parts) >= 3 else chrom
    rng = parts[1].split("-")
    start = int(rng[0].replace(",", ""))
    end = int(rng[1].replace(",", "")) if len(rng) >= 2 else start
    start -= extend
    end += extend
    if zero_based and start < end:
        start += 1
    if start > end:
        start = end
    return [[(chrom, start, end, gene)]]


def from_bed(lines, names, given, extend=0, zero_based=False, delimiter="\t"):
    """Launcher::readBedFile (:134-184) and RegionBuilder::buildRegions (:45-104): one list of
    regions per row (its segments)."""
    col = columns(given)
    segs = []
    for line in lines:
        line = line.rstrip("\n")
        if line.startswith("#") or line.startswith("browser") or line.startswith("track"):
            continue
        v = line.split(delimiter)
        chrom = correct_chromosome(v[col["c"]], names)
        cds_start, cds_end = int(v[col["S"]]), int(v[col["E"]])
        gene = v[col["g"]] if col["g"] < len(v) else chrom
        starts, ends = v[col["s"]].split(","), v[col["e"]].split(",")
        row = []
        for i in range(len(starts)):
            if starts[i] == "":
                continue
            ts, te = int(starts[i]), int(ends[i])
            if cds_start > te:
                continue
            if cds_end > te:
                break
            ts = max(ts, cds_start)
            te = min(te, cds_end)
            ts -= extend
            te += extend
            if zero_based and ts < te:
                ts += 1
            row.append((chrom, ts, te, gene))
        segs.append(row)
    return segs


def auto_resize(segs):
    """RegionBuilder::AdjustRegionSize (:15-37): a row whose first region is longer than 10 kb becomes
    10 kb pieces of that region (its other segments are dropped, as the reference does)."""
    out = []
    for row in segs:
        if not row:
            out.append(row)
            continue
        chrom, start, end, gene = row[0]
        if end - start > REGION_SIZE_MAX:
            while end - start > REGION_SIZE_MAX:
                out.append([(chrom, start, start + REGION_SIZE_MAX, gene)])
                start += REGION_SIZE_MAX
            out.append([(chrom, start, end, gene)])
        else:
            out.append(row)
    return out


def flatten(segs):
    return [r for row in segs for r in row]
