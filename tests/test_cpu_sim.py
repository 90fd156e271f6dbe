"""The engine's own stage code, compiled for the host (tests/cpu_sim: device bodies run one item at a
time with host stand-ins for the atomics), against the oracle.  Catches logic regressions in the
HIP path without a GPU; the GPU tests run the same comparisons on the device."""
import os
import random
import subprocess

import pytest

from conftest import ROOT, run_oracle

SIM_DIR = os.path.join(ROOT, "tests", "cpu_sim")
SIM = os.path.join(SIM_DIR, "_build", "rv_cpu_sim")
CSRC = os.path.join(ROOT, "rabbitvar_amd", "csrc")


@pytest.fixture(scope="session")
def sim():
    os.makedirs(os.path.dirname(SIM), exist_ok=True)
    srcs = [os.path.join(SIM_DIR, "rv_cpu_sim.cpp")] + [os.path.join(CSRC, f) for f in
                                                        ("rv_host.cpp", "rv_realign.cpp", "rv_format.cpp",
                                                         "rv_order.cpp")]
    tmp = f"{SIM}.{os.getpid()}"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", SIM_DIR, "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-o",
                    tmp] + srcs, check=True)
    os.replace(tmp, SIM)  # atomic: parallel sessions never run a half-written binary
    return SIM


@pytest.mark.parametrize("case,k,fisher", [("c1", 1, 1), ("edge", 1, 1), ("edge", 0, 0)])
def test_engine_stages_match_oracle(sim, dataset, case, k, fisher):
    d = dataset(case)
    with open(os.path.join(d, "ref.fa")) as f:
        L = sum(len(x.strip()) for x in f if not x.startswith(">"))
    extra = ["-k", str(k)] + (["--fisher"] if fisher else [])
    want, _ = run_oracle(d, f"chrS:1-{L}", extra, out=f"sim_{k}_{fisher}.txt")
    out = os.path.join(d, f"cpu_sim_{k}_{fisher}.txt")
    subprocess.run([sim, os.path.join(d, "ref.fa"), os.path.join(d, "reads.rvb"), "1", str(L), str(k), str(fisher), out],
                   check=True)
    with open(out) as f:
        got = f.read()
    assert len(want.splitlines()) > 100
    assert sorted(got.splitlines()) == sorted(want.splitlines())
    assert got == want  # same lines in the same (robin_hood slot) order


@pytest.mark.parametrize("samfilter", ["504", "0"])
def test_engine_stages_duplicate_filter(sim, dataset, samfilter):
    """-t through the device duplicate-filter body, with the default -F and with -F 0 (which lets the
    reads flagged unmapped reach the CIGAR-keyed branch)."""
    d = dataset("dups")
    extra = ["--fisher", "-t", "-F", samfilter]
    want, _ = run_oracle(d, "chrS:1-60000", extra, out=f"sim_dup_{samfilter}.txt")
    plain, _ = run_oracle(d, "chrS:1-60000", ["--fisher", "-F", samfilter], out=f"sim_nodup_{samfilter}.txt")
    out = os.path.join(d, f"cpu_sim_dup_{samfilter}.txt")
    env = dict(os.environ, RV_SIM_DEDUP="1", RV_SIM_SAMFILTER=samfilter)
    subprocess.run([sim, os.path.join(d, "ref.fa"), os.path.join(d, "reads.rvb"), "1", "60000", "1", "1", out],
                   check=True, env=env, capture_output=True)
    with open(out) as f:
        got = f.read()
    assert len(want.splitlines()) > 500 and want != plain  # the filter changes the calls
    assert got == want


@pytest.mark.parametrize("case,start,end", [("c1", 1, 20000), ("edge", 29001, 31000)])
def test_engine_stages_pileup_mode(sim, dataset, case, start, end):
    """-p: every covered position prints, reference-only ones in their reference-variant form."""
    d = dataset(case)
    want, _ = run_oracle(d, f"chrS:{start}-{end}", ["--fisher", "-p"], out=f"sim_pileup_{start}.txt")
    out = os.path.join(d, f"cpu_sim_pileup_{start}.txt")
    env = dict(os.environ, RV_SIM_PILEUP="1")
    subprocess.run([sim, os.path.join(d, "ref.fa"), os.path.join(d, "reads.rvb"), str(start), str(end), "1", "1", out],
                   check=True, env=env, capture_output=True)
    with open(out) as f:
        got = f.read()
    assert len(want.splitlines()) > 0.9 * (end - start + 1)
    assert got == want


@pytest.mark.parametrize("start,end", [(1, 1), (5000, 5150), (59000, 60000)])
def test_engine_stages_subregions(sim, dataset, start, end):
    d = dataset("edge")
    want, _ = run_oracle(d, f"chrS:{start}-{end}", ["--fisher"], out=f"simsub_{start}.txt")
    out = os.path.join(d, f"cpu_sim_sub_{start}.txt")
    subprocess.run([sim, os.path.join(d, "ref.fa"), os.path.join(d, "reads.rvb"), str(start), str(end), "1", "1", out],
                   check=True)
    with open(out) as f:
        got = f.read()
    assert sorted(got.splitlines()) == sorted(want.splitlines())
    assert got == want


def test_char_map_order_matches_oracle_clone(sim, tools):
    """The engine's replay of robin_hood<char,int> iteration order (soft-clip consensus ties) against
    the oracle's full clone, for random insertion sequences of up to 16 distinct bytes."""
    probe = os.path.join(ROOT, "oracle", "_build", "rhmap_probe")
    if not os.path.exists(probe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    rnd = random.Random(3)
    alphabet = [ord(c) for c in "ACGTN"] + list(range(0, 256, 17))
    for _ in range(200):
        k = rnd.randint(1, 16)
        seq = rnd.sample(sorted(set(alphabet)), min(k, len(set(alphabet))))
        r = subprocess.run([probe], input="".join(f"c {b}\n" for b in seq), capture_output=True, text=True, check=True)
        want = [int(x) for x in r.stdout.splitlines()[-1].split()[1:]]
        g = subprocess.run([sim, "--rh-chars"] + [str(b) for b in seq], capture_output=True, text=True, check=True)
        got = [int(x) for x in g.stdout.split()]
        assert got == want, seq


@pytest.mark.parametrize("seed", range(6))
def test_int_map_order_matches_oracle_clone(sim, tools, seed):
    """The engine's robin_hood slot-order model (rv_rhorder.h) against the oracle's clone, on insert /
    erase sequences large enough to grow the table many times and to halve the info increment."""
    probe = os.path.join(ROOT, "oracle", "_build", "rhmap_probe")
    if not os.path.exists(probe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    rnd = random.Random(seed)
    n = [100, 5000, 60000, 200000, 20000, 1000][seed]
    ops, live = [], []
    for i in range(n):
        if live and rnd.random() < 0.1:
            k = live.pop(rnd.randrange(len(live)))
            ops.append(f"e {k}\n")
        else:
            k = rnd.randrange(-1000, 10**7) if seed % 2 else i * 3 + rnd.randrange(3)
            live.append(k)
            ops.append(f"i {k}\n")
    text = "".join(ops)
    r = subprocess.run([probe], input=text, capture_output=True, text=True, check=True)
    want = [int(ln.split()[1]) for ln in r.stdout.splitlines()[1:-1]]
    g = subprocess.run([sim, "--rh-ints"], input=text, capture_output=True, text=True, check=True)
    got = [int(x) for x in g.stdout.split()]
    assert got == want


@pytest.mark.parametrize("n,seed", [(3400, 1), (27000, 2), (1700000, 5)])
def test_sorted_slot_order_with_realigner_log(sim, n, seed):
    """The sorted form of nonInsertionVariants' slot order, with random realigner insertions and
    erasures, against the full replay: whenever it claims to apply it must be identical."""
    r = subprocess.run([sim, "--order-check", str(n), str(seed)], capture_output=True, text=True, check=True)
    applied, equal = (int(x) for x in r.stdout.split())
    assert applied == equal and applied > 0


@pytest.mark.parametrize("n,seed", [(2000, 1), (60000, 2), (400000, 3), (2500000, 4)])
def test_per_growth_sorted_order_matches_replay(sim, n, seed):
    """alignedVariants' slot order by the per-growth sort (growth epochs, info-increment halvings
    found by simulating only the runs that can reach them) against the full robin_hood replay, on
    insert and insert-then-erase sequences with crowded home ranges: identical whenever it claims
    to apply, and the sequences reach halved increments."""
    r = subprocess.run([sim, "--aligned-check", str(n), str(seed)], capture_output=True, text=True, check=True)
    applied, equal, halved, bailed, min_inc = (int(x) for x in r.stdout.split())
    print(f"n={n}: applied {applied} equal {equal} halved {halved} bailed {bailed} smallest increment {min_inc}")
    assert applied == equal and applied > 0
    if n >= 60000:
        assert halved > 0 and min_inc <= 8  # two or more halvings in a sequence the sort applied to


def test_find35match_against_nested_loops(sim):
    """find35match's per-diagonal evaluation against the reference's nested loops
    (src/VariationRealigner.cpp:1396-1428) restated directly, on random clip consensus pairs of
    11-160 bases, half of them built to overlap: the same (b5, b3, score) for every pair."""
    r = subprocess.run([sim, "--find35-check", "6000", "11"], capture_output=True, text=True, check=True)
    trials, matched, equal = (int(x) for x in r.stdout.split())
    assert equal == trials and matched > 100


def test_fixed_point_formatting_matches_printf(sim):
    """The lines' "%f" fields come from an integer formatter (rv_format.cpp put_f); it must print
    exactly what printf prints, ties to even included."""
    r = subprocess.run([sim, "--fmt-check", "400000", "7"], capture_output=True, text=True, check=True)
    assert r.stdout.strip() == "0", r.stderr


@pytest.fixture(scope="session")
def sim_asan():
    """The same harness built with AddressSanitizer (host code only)."""
    out = os.path.join(SIM_DIR, "_build", "rv_cpu_sim_asan")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = [os.path.join(SIM_DIR, "rv_cpu_sim.cpp")] + [os.path.join(CSRC, f) for f in
                                                        ("rv_host.cpp", "rv_realign.cpp", "rv_format.cpp",
                                                         "rv_order.cpp")]
    tmp = f"{out}.{os.getpid()}"
    subprocess.run(["g++", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-std=c++17", "-I", SIM_DIR,
                    "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs, check=True)
    os.replace(tmp, out)
    return out


def _long_cigar_records(ref, long_at, n=12, start=200, rlen=100):
    """n reads at start+1; those at indexes `long_at` carry 59 CIGAR operations (29 x 2M1D, then 42M)."""
    from rabbitvar_amd import batch_from_records

    recs = []
    for i in range(n):
        if i in long_at:
            cig = [(0, 2), (2, 1)] * 29 + [(0, rlen - 58)]
        else:
            cig = [(0, rlen)]
        seq, rp = [], start
        for op, ln in cig:
            if op == 0:
                seq.append(ref[rp:rp + ln])
                rp += ln
            else:
                rp += ln
        seq = "".join(seq)
        nm = sum(ln for op, ln in cig if op == 2)
        recs.append({"pos0": start, "flag": 16 if i % 2 else 0, "mapq": 60, "nm": nm, "seq": seq,
                     "qual": [30] * len(seq), "cigar": cig})
    return batch_from_records(recs)


@pytest.mark.parametrize("long_at", [(5,), (11,), (5, 11)])
def test_long_cigar_reads_are_refused(sim_asan, tmp_path, long_at):
    """A read whose CIGAR has more operations than its 48-word record slot allows (here 59, in the
    middle of the batch and as its last read) is refused before the slot is written, and the
    region fails with a capacity error (exit 4) instead of walking a cut record.  Under ASan no
    access leaves the record buffer, and every other read's prep record (status, position,
    rewritten CIGAR words) equals the one it gets in a batch without the long reads."""
    import handbuilt as hb
    from rabbitvar_amd import write_rvb

    ref = hb.reference()
    fa = str(tmp_path / "ref.fa")
    with open(fa, "w") as f:
        f.write(">chrS\n" + ref + "\n")
    with_long, without = str(tmp_path / "a.rvb"), str(tmp_path / "b.rvb")
    write_rvb(with_long, _long_cigar_records(ref, long_at))
    write_rvb(without, _long_cigar_records(ref, (), n=12 - len(long_at)))
    da, db = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    ra = subprocess.run([sim_asan, fa, with_long, "1", "600", "1", "1", str(tmp_path / "oa.txt")],
                        env=dict(os.environ, RV_SIM_PREP_DUMP=da), capture_output=True, text=True)
    assert ra.returncode == 4, ra.stderr
    assert f"capacity: {len(long_at)} read(s) refused" in ra.stderr
    assert "AddressSanitizer" not in ra.stderr
    rb = subprocess.run([sim_asan, fa, without, "1", "600", "1", "1", str(tmp_path / "ob.txt")],
                        env=dict(os.environ, RV_SIM_PREP_DUMP=db), capture_output=True, text=True)
    assert rb.returncode == 0, rb.stderr
    a = open(da).read().splitlines()
    b = open(db).read().splitlines()
    kept = [ln for i, ln in enumerate(a) if i not in long_at]
    assert kept == b
    for i in long_at:
        assert a[i].split()[0] == "3"  # RD_NOWALK: never walked


def test_cigar_rewrite_known_answers_engine(sim, tmp_path):
    """The engine's CIGAR rewrite (rv_cigar.h through prep_one, as k_prep runs it) on the hand-worked
    known-answer reads of test_oracle_cigar_kat.py: the position and the CIGAR it leaves for the walk
    equal the hand-derived answers."""
    import test_oracle_cigar_kat as kat
    from rabbitvar_amd import batch_from_records, write_rvb

    names = list(kat.CASES)
    rows = [kat.CASES[n]() for n in names]
    recs = []
    for P, cig, seq, q, _ in rows:
        ops, num = [], ""
        for ch in cig:
            if ch.isdigit():
                num += ch
            else:
                ops.append(("MIDNSHP=X".index(ch), int(num)))
                num = ""
        quals = [int(x) for x in q.split(",")] if isinstance(q, str) else [q] * len(seq)
        idlen = sum(n for o, n in ops if o in (1, 2))
        recs.append({"pos0": P - 1, "flag": 0, "mapq": 60, "nm": idlen, "seq": seq, "qual": quals, "cigar": ops})
    fa = str(tmp_path / "ref.fa")
    with open(fa, "w") as f:
        f.write(">chrS\n" + kat.REF + "\n")
    rvb = str(tmp_path / "kat.rvb")
    write_rvb(rvb, batch_from_records(recs))
    dump = str(tmp_path / "prep.txt")
    subprocess.run([sim, fa, rvb, "1", "600", "1", "1", str(tmp_path / "o.txt")], check=True, capture_output=True,
                   env=dict(os.environ, RV_SIM_PREP_DUMP=dump))
    # the harness keeps the region's records in iterator order (sorted input is not required here:
    # every KAT read overlaps the region)
    lines = open(dump).read().splitlines()
    assert len(lines) == len(rows)
    bad = []
    for name, (P, cig, seq, q, want), ln in zip(names, rows, lines):
        f = [int(x) for x in ln.split()]
        status, pos, n = f[0], f[1], f[2]
        text = "".join(f"{w >> 4}{'MIDNSHP=XB'[w & 15]}" for w in f[3:3 + n])
        if status != 0 or (pos, text) != want[1:]:
            bad.append((name, P, cig, status, pos, text, want))
    assert not bad, bad


def test_descriptor_matchers_equal_std_regex(sim):
    """The realigner's hand-written descriptor matchers (rv_desc.h) give std::regex's match and
    captures, with the reference's pattern texts (include/patterns.h), on 200,000 random strings over
    the descriptor alphabet and descriptor-shaped pieces."""
    r = subprocess.run([sim, "--desc-fuzz", "200000", "5"], capture_output=True, text=True, check=True)
    trials, bad, hits = (int(x) for x in r.stdout.split())
    assert bad == 0, r.stderr
    assert hits > 100000


@pytest.mark.parametrize("case", ["mnp", "realigndel", "realignins", "lgins30", "caret", "amp"])
def test_realigner_known_answers_engine(sim, tmp_path, case):
    """The realigner and deletion-descriptor known-answer reads through the engine's device stages
    and host realigner on the CPU: the same lines as the oracle, whose answers
    tests/test_oracle_realign_kat.py and tests/test_oracle_allele_kat.py pin."""
    import test_oracle_allele_kat as ak
    import test_oracle_realign_kat as rk

    d = str(tmp_path / case)
    {"mnp": rk.build, "realigndel": rk.build_realigndel, "realignins": rk.build_realignins,
     "lgins30": rk.build_lgins30, "caret": lambda x: ak.build(x, "caret"), "amp": lambda x: ak.build(x, "amp")}[case](d)
    want, _ = run_oracle(d, "chrS:1-600", ["--fisher"])
    out = os.path.join(d, "cpu_sim.txt")
    subprocess.run([sim, os.path.join(d, "ref.fa"), os.path.join(d, "reads.rvb"), "1", "600", "1", "1", out],
                   check=True, capture_output=True)
    with open(out) as f:
        got = f.read()
    assert got == want and len(want.splitlines()) >= 1
