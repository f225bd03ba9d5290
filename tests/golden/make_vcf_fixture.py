"""Write tests/golden/test200.npz from the reference's end-to-end test data.

Inputs (reference repository, read as data):
  test/endtoend/input/test-200-samples.vcf.gz        200 diploid individuals, phased
  test/endtoend/expect/test-200-samples.freq.txt     ALT COUNT per (position, ref, alt)
  test/endtoend/expect/test-200-samples.zygosity.txt AA / Aa / aa per (position, ref, alt)

The expected files are what the reference's tests compare `grg process freq`,
`dot_product(ones, UP)` (test/endtoend/test_basic.py:71-96) and the zygosity report
(:134-149) against.  Allele counts are the UP product with all-ones input on ANY correct
GRG of this VCF, so they pin the UP traversal independently of graph construction.

Output arrays (one mutation per (record, alt allele), in VCF order, which is the
(position, ref, alt) order of the expect files):
  positions, carrier_off, carriers   haplotype indices carrying each alt allele
  expect_alt_count, expect_AA, expect_Aa, expect_aa
  num_samples (haplotypes), ploidy
"""

import gzip
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    vcf = os.path.join(REF, "test/endtoend/input/test-200-samples.vcf.gz")
    freq = os.path.join(REF, "test/endtoend/expect/test-200-samples.freq.txt")
    zyg = os.path.join(REF, "test/endtoend/expect/test-200-samples.zygosity.txt")
    keys, carriers, positions = [], [], []
    num_samples = None
    with gzip.open(vcf, "rt") as f:
        for line in f:
            if line.startswith("##"):
                continue
            cols = line.rstrip("\n").split("\t")
            if line.startswith("#"):
                num_samples = 2 * (len(cols) - 9)
                continue
            pos, ref, alts = int(cols[1]), cols[3], cols[4].split(",")
            gts = []
            for gt in cols[9:]:
                a, b = gt.replace("/", "|").split("|")
                gts.append(a)
                gts.append(b)
            for ai, alt in enumerate(alts, start=1):
                c = [h for h, g in enumerate(gts) if g == str(ai)]
                keys.append((pos, ref, alt))
                carriers.append(c)
                positions.append(pos)
    expect = {}
    with open(freq) as f:
        for line in f:
            p = line.strip().split("\t")
            if not p or p[0] == "POSITION":
                continue
            expect[(int(float(p[0])), p[1], p[2])] = int(p[3])
    zexp = {}
    with open(zyg) as f:
        for line in f:
            p = line.strip().split("\t")
            if not p or p[0] == "POSITION":
                continue
            zexp[(int(float(p[0])), p[1], p[2])] = (int(p[3]), int(p[4]), int(p[5]))
    assert len(keys) == len(expect) == len(zexp), (len(keys), len(expect), len(zexp))
    off = np.zeros(len(carriers) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(c) for c in carriers])
    flat = np.array([h for c in carriers for h in c], dtype=np.uint32)
    np.savez_compressed(
        os.path.join(HERE, "test200.npz"),
        positions=np.array(positions, dtype=np.int64),
        carrier_off=off,
        carriers=flat,
        expect_alt_count=np.array([expect[k] for k in keys], dtype=np.int64),
        expect_AA=np.array([zexp[k][0] for k in keys], dtype=np.int64),
        expect_Aa=np.array([zexp[k][1] for k in keys], dtype=np.int64),
        expect_aa=np.array([zexp[k][2] for k in keys], dtype=np.int64),
        num_samples=np.int64(num_samples),
        ploidy=np.int64(2),
    )
    print("mutations", len(keys), "haplotypes", num_samples, "carrier entries", len(flat))


if __name__ == "__main__":
    main()
