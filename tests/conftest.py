import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SYNTH = os.path.join(ROOT, "rabbitvar_amd", "_build", "rvsynth")
ORACLE = os.path.join(ROOT, "oracle", "_build", "rvcpu")

# Synthetic workloads used by the parity tests (small enough for the oracle to finish in seconds).
CASES = {
    # 200 kb contig, 30x, SNVs 1e-3, indels 1-10 bp at 1e-4
    "c1": ["--seed", "1", "--len", "200000", "--depth", "30"],
    # 300 kb, 60x, 2% tandem-repeat tracts, indels 1-30 bp at 1/8000, 12 long deletions (clipped reads)
    "c2s": ["--seed", "2", "--len", "300000", "--depth", "60", "--indel-rate", "0.000125", "--indel-max", "30",
            "--long-dels", "12", "--repeat-frac", "0.02"],
    # dense variant edge cases: high indel rate, long indels, low quality tail errors
    "edge": ["--seed", "5", "--len", "60000", "--depth", "40", "--indel-rate", "0.002", "--indel-max", "40",
             "--snv-rate", "0.004", "--long-dels", "6", "--repeat-frac", "0.05", "--err-rate", "0.01"],
    # duplicate-filter shapes (-t): 30% of records duplicated, half unpaired, 5% flagged unmapped
    "dups": ["--seed", "11", "--len", "60000", "--depth", "40", "--indel-rate", "0.001", "--dup-rate", "0.3",
             "--single-end-frac", "0.5", "--unmapped-frac", "0.05"],
    # tumor/normal pair: 200 kb, tumor 60x, normal 40x, somatic SNVs/indels at 5e-4 with VAF 0.05-0.4
    "som": ["--seed", "5", "--len", "200000", "--depth", "60", "--normal-depth", "40", "--somatic-rate", "0.0005",
            "--vaf", "0.05", "0.4", "--indel-rate", "0.0002"],
    # deep tumor/normal pair: 40 kb, tumor 250x, normal 120x, somatic variants at VAF 0.01-0.2, 2% duplicates
    "deep": ["--seed", "13", "--len", "40000", "--depth", "250", "--normal-depth", "120", "--somatic-rate", "0.001",
             "--vaf", "0.01", "0.2", "--indel-rate", "0.0005", "--dup-rate", "0.02"],
    # the 5 Mb, 60x germline workload at its full size (the bench workload's shape, another seed)
    "c2full": ["--seed", "8", "--len", "5000000", "--depth", "60", "--indel-rate", "0.000125", "--indel-max", "30",
               "--long-dels", "200", "--repeat-frac", "0.02"],
    # the deep somatic pair at full depth on a 1 Mb slice: tumor 200x, normal 100x, somatic SNVs and
    # indels at 6e-5/bp with VAF 0.01-0.2, 2% duplicated records
    "c3slice": ["--seed", "31", "--len", "1000000", "--depth", "200", "--normal-depth", "100", "--somatic-rate",
                "0.00006", "--vaf", "0.01", "0.2", "--indel-rate", "0.0001", "--dup-rate", "0.02"],
    # tumor/normal pair with the dense edge-case mix above and a thin normal
    "som_edge": ["--seed", "7", "--len", "60000", "--depth", "40", "--normal-depth", "30", "--indel-rate", "0.002",
                 "--indel-max", "40", "--snv-rate", "0.004", "--long-dels", "6", "--repeat-frac", "0.05",
                 "--err-rate", "0.01", "--somatic-rate", "0.001"],
}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_tools():
    if not os.path.exists(ORACLE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(SYNTH):
        os.makedirs(os.path.dirname(SYNTH), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", SYNTH,
                        os.path.join(ROOT, "rabbitvar_amd", "synth", "rv_synth.cpp"), "-lz"], check=True)


@pytest.fixture(scope="session")
def tools():
    _ensure_tools()
    return {"oracle": ORACLE, "synth": SYNTH}


@pytest.fixture(scope="session")
def dataset(tmp_path_factory, tools):
    cache = {}

    def get(name):
        if name not in cache:
            d = tmp_path_factory.mktemp(name)
            subprocess.run([tools["synth"], "--out", str(d)] + CASES[name], check=True, capture_output=True)
            cache[name] = str(d)
        return cache[name]

    return get


def run_oracle(d, region, extra=(), dump=None, out="out.txt", reads="reads.rvb"):
    cmd = [ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", os.path.join(d, reads), "-N", "S1", "-R", region,
           "--out", os.path.join(d, out)] + list(extra)
    if dump:
        cmd += ["--dump-pileup", os.path.join(d, dump)]
    subprocess.run(cmd, check=True, capture_output=True)
    with open(os.path.join(d, out)) as f:
        text = f.read()
    dump_text = None
    if dump:
        with open(os.path.join(d, dump)) as f:
            dump_text = f.read()
    return text, dump_text


def run_oracle_somatic(d, region, extra=(), out="somatic.txt"):
    """The oracle in somatic mode (-b 'tumor|normal' -N 'T|N') over one region."""
    bams = os.path.join(d, "reads.rvb") + "|" + os.path.join(d, "normal.rvb")
    cmd = [ORACLE, "-G", os.path.join(d, "ref.fa"), "-b", bams, "-N", "T|N", "-R", region,
           "--out", os.path.join(d, out)] + list(extra)
    subprocess.run(cmd, check=True, capture_output=True)
    with open(os.path.join(d, out)) as f:
        return f.read()
