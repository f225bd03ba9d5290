"""Write tests/golden/test200_miss.npz from the reference's end-to-end test data.

Input (reference repository, read as data):
  test/endtoend/input/test-200-samples.miss.igd   the 200-individual test VCF with some genotypes
                                                  set missing, in IGD format (picovcf's IGDData:
                                                  a 128-byte header, an index of 16-byte entries
                                                  (position with sparse / missing flags, file
                                                  offset), rows stored as a u32 count and u32
                                                  sample IDs when sparse, else a bit vector,
                                                  most significant bit first)

The file holds the 10,893 variants of test/endtoend/input/test-200-samples.vcf.gz (the rows of
tests/golden/test200.npz) with the missing samples removed from each variant at a missing
site, plus 18 missing-data rows.  This script checks that, and stores only the missing rows:
  miss_positions, miss_off, miss_carriers   position and missing haplotypes of each row
The reference's test (test/endtoend/test_missing.py:19-53) states the two totals a correct GRG
of this file gives: 25 missing samples and 21 missing individuals, summed over the mutations
that have both a node and a missingness node.
"""

import os
import struct
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

SPARSE = 0x0100000000000000
MISSING = 0x0200000000000000
POS_MASK = 0x0000FFFFFFFFFFFF


def read_igd(path):
    b = open(path, "rb").read()
    magic, version, ploidy, _, num_var, num_ind, _, flags, pos_index = struct.unpack_from("<QQIIQIIQQ", b, 0)
    assert magic == 0x3A0C6FD7945A3481 and version in (3, 4) and flags & 1, "expected a phased IGD file"
    n = num_ind * ploidy
    rows = []
    for v in range(num_var):
        bp, fp = struct.unpack_from("<QQ", b, pos_index + 16 * v)
        assert (bp >> 48) & 0xFF == 0, "unphased copy counts are not expected here"
        if bp & SPARSE:
            (cnt,) = struct.unpack_from("<I", b, fp)
            s = np.frombuffer(b, dtype="<u4", count=cnt, offset=fp + 4).astype(np.uint32)
        else:
            bits = np.unpackbits(np.frombuffer(b, dtype=np.uint8, count=(n + 7) // 8, offset=fp))
            s = np.nonzero(bits[:n])[0].astype(np.uint32)
        rows.append((bp & POS_MASK, bool(bp & MISSING), np.sort(s)))
    return n, ploidy, rows


def main():
    n, ploidy, rows = read_igd(os.path.join(REF, "test/endtoend/input/test-200-samples.miss.igd"))
    fx = np.load(os.path.join(HERE, "test200.npz"))
    assert n == int(fx["num_samples"]) and ploidy == int(fx["ploidy"])
    miss = {}
    for pos, is_miss, s in rows:
        if is_miss:
            miss.setdefault(int(pos), set()).update(s.tolist())
    plain = [r for r in rows if not r[1]]
    off, car, positions = fx["carrier_off"], fx["carriers"], fx["positions"]
    assert len(plain) == len(positions)
    for i, (pos, _, s) in enumerate(plain):
        assert pos == positions[i]
        want = sorted(set(car[off[i]:off[i + 1]].tolist()) - miss.get(int(pos), set()))
        assert s.tolist() == want, i
    mpos = sorted(miss)
    moff = np.zeros(len(mpos) + 1, dtype=np.uint64)
    moff[1:] = np.cumsum([len(miss[p]) for p in mpos])
    mcar = np.array([x for p in mpos for x in sorted(miss[p])], dtype=np.uint32)
    np.savez_compressed(os.path.join(HERE, "test200_miss.npz"), miss_positions=np.array(mpos, dtype=np.int64),
                        miss_off=moff, miss_carriers=mcar)
    print(f"{len(mpos)} missing rows, {len(mcar)} missing haplotypes")


if __name__ == "__main__":
    main()
