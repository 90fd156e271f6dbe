"""Region planning (rabbitvar_amd/regions.py) on cases worked out by hand from the reference's rules
(src/RegionBuilder.cpp:15-217, src/Launcher.cpp:21, :134-184, :415-430)."""
from rabbitvar_amd import regions as rg

NAMES = {"chr1", "chrS", "2"}


def test_region_option():
    assert rg.from_option("chr1:1,000-2,000", NAMES) == [[("chr1", 1000, 2000, "chr1")]]
    assert rg.from_option("1:100-200:BRCA", NAMES) == [[("chr1", 100, 200, "BRCA")]]  # "chr" added
    assert rg.from_option("chr2:5-9", NAMES) == [[("2", 5, 9, "2")]]                    # "chr" dropped
    assert rg.from_option("chrS:500", NAMES) == [[("chrS", 500, 500, "chrS")]]
    assert rg.from_option("chrS:100-200", NAMES, extend=10) == [[("chrS", 90, 210, "chrS")]]
    assert rg.from_option("chrS:100-200", NAMES, zero_based=True) == [[("chrS", 101, 200, "chrS")]]
    assert rg.from_option("chrS:300-200", NAMES) == [[("chrS", 200, 200, "chrS")]]


def test_columns():
    assert rg.columns({}) == {"c": 2, "S": 6, "E": 7, "s": 9, "e": 10, "g": 12}
    assert rg.columns({"c": 1, "S": 2, "E": 3, "g": 4}) == {"c": 0, "S": 1, "E": 2, "s": 1, "e": 2, "g": 3}
    assert rg.columns({"c": 1, "S": 2, "E": 3, "s": 5, "e": 6}) == {"c": 0, "S": 1, "E": 2, "s": 4, "e": 5, "g": 12}


def test_bed_rows():
    given = {"c": 1, "S": 2, "E": 3, "g": 4}
    lines = ["track name=x\n", "#c\n", "browser y\n", "chrS\t100\t200\tA\n", "S\t300\t400\n"]
    assert rg.flatten(rg.from_bed(lines, NAMES, given)) == [("chrS", 100, 200, "A"), ("chrS", 300, 400, "chrS")]
    assert rg.flatten(rg.from_bed(lines, NAMES, given, extend=5, zero_based=True)) == [
        ("chrS", 96, 205, "A"), ("chrS", 296, 405, "chrS")]


def test_refgene_segments():
    # refGene-style row: cds 1000-2000.  The reference stops at the first segment that ends before
    # the cds end (RegionBuilder.cpp:88-90 tests cdsEnd > thickEnd), so a row whose first exon ends
    # inside the cds gives no region at all, and segments are clamped to the cds otherwise.
    def row(starts, ends):
        return "\t".join(["0", "NM_1", "chr1", "+", "800", "2700", "1000", "2000", "2", starts, ends, "0",
                          "GENE1"]) + "\n"

    assert rg.from_bed([row("900,1500,", "1100,1600,")], NAMES, {}) == [[]]
    assert rg.from_bed([row("500,900,", "800,2100,")], NAMES, {}) == [[("chr1", 1000, 2000, "GENE1")]]
    assert rg.from_bed([row("900,2500,", "2100,2600,")], NAMES, {}, zero_based=True) == [
        [("chr1", 1001, 2000, "GENE1"), ("chr1", 2500, 2000, "GENE1")]]  # starts are not clamped to the end


def test_auto_resize():
    segs = [[("chrS", 1, 25000, "g")], [("chrS", 30001, 30500, "h"), ("chrS", 31000, 31100, "h")]]
    assert rg.auto_resize(segs) == [[("chrS", 1, 10001, "g")], [("chrS", 10001, 20001, "g")],
                                    [("chrS", 20001, 25000, "g")],
                                    [("chrS", 30001, 30500, "h"), ("chrS", 31000, 31100, "h")]]
