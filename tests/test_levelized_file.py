"""Levelized-graph files on the host (grgl_hip_levelize_to_file, grgl_hip_levelized_file_info;
the file half of the reference's storeGPUGRGToDisk / loadGPUGRGFromDisk,
src/cuda/gpu_grg.cu:10-141): sizes reported from the file, and every corruption rejected, both
the ones the checksum catches and structurally invalid plans with a correct checksum (which
would otherwise make a traversal read out of bounds or out of order)."""

import struct

import numpy as np
import pytest

import grgl_amd
from grgl_amd import capi
from helpers import random_grg

M64 = (1 << 64) - 1


def _mix(h, w):
    h ^= (w + 0x9E3779B97F4A7C15 + ((h << 6) & M64) + (h >> 2)) & M64
    h = (h * 0xBF58476D1CE4E5B9) & M64
    return h ^ (h >> 31)


def _checksum(data):
    """The file's checksum: 8-byte little-endian words (tail zero-padded) through a
    multiply-xorshift mix, then a final constant."""
    h = 0x243F6A8885A308D3
    n = len(data) // 8 * 8
    for w in np.frombuffer(data[:n], dtype="<u8").tolist():
        h = _mix(h, w)
    if n < len(data):
        h = _mix(h, int.from_bytes(data[n:].ljust(8, b"\0"), "little"))
    return _mix(h, 0x5BD1E995)


ARRAYS = ["levelStart", "order", "rank", "off", "src", "outRow", "missRow", "coalByRow"]


def _parse(buf):
    """Byte offsets of the fields: header, then per plan its scalars and arrays."""
    assert buf[:8] == b"GRGLLVLZ"
    pos = 8 + 4 + 5 * 4 + 8
    plans = []
    for _ in range(2):
        p = {"scalars": pos}
        pos += 4 * 4 + 8
        for name in ARRAYS:
            (n,) = struct.unpack_from("<Q", buf, pos)
            p[name] = (pos + 8, n)
            pos += 8 + 4 * n
        plans.append(p)
    assert pos + 8 == len(buf)
    return plans


def _array(buf, field):
    off, n = field
    return np.frombuffer(bytes(buf[off:off + 4 * n]), dtype="<u4").copy()


def _rewrite(path, buf):
    body = bytes(buf[:-8])
    with open(path, "wb") as f:
        f.write(body + struct.pack("<Q", _checksum(body)))


@pytest.fixture(scope="module")
def rich():
    return random_grg(9, num_samples=60, num_internal=300, max_children=7, heavy=(80, 200), num_mutations=400,
                      with_miss=True, with_coal=True)


def test_checksum_restatement_matches(tmp_path, rich):
    path = str(tmp_path / "g.lvlz")
    capi.levelize_to_file(rich, path)
    buf = open(path, "rb").read()
    assert struct.unpack("<Q", buf[-8:])[0] == _checksum(buf[:-8])


@pytest.mark.parametrize("which", ["rich", "coalescent", "constructed"])
def test_file_info_matches_graph(tmp_path, rich, which):
    if which == "rich":
        arrays = rich
    elif which == "coalescent":
        arrays = grgl_amd.grg_from_coalescent(500, 4000, 3, seed=2).arrays()
    else:
        src = grgl_amd.grg_from_coalescent(400, 3000, 2, seed=4)
        off, car = src.mutation_carriers()
        arrays = grgl_amd.fast_grg_from_samples(400, 2, carrier_offsets=off, carriers=car).arrays()
    path = str(tmp_path / "g.lvlz")
    capi.levelize_to_file(arrays, path)
    info = capi.levelized_file_info(path)
    V = len(arrays["down_offsets"]) - 1
    assert (info["num_samples"], info["num_nodes"], info["num_mutations"], info["ploidy"]) == (
        int(arrays["num_samples"]), V, len(arrays["mutation_nodes"]), int(arrays["ploidy"]))
    assert info["num_edges"] == int(arrays["down_offsets"][-1]) == info["up_gather_entries"]
    miss = arrays.get("mutation_miss_nodes")
    seeds = int(np.sum(arrays["mutation_nodes"] < V)) + (0 if miss is None else int(np.sum(miss < V)))
    assert info["down_gather_entries"] == info["num_edges"] + seeds
    g = grgl_amd.grg_from_arrays(arrays["num_samples"], arrays["ploidy"], arrays["down_offsets"],
                                 arrays["down_children"], arrays["mutation_nodes"])
    assert info["up_levels"] == g.graph_stats()["up_levels"]
    assert info["device"] == -1 and info["device_bytes"] > 0


def test_damaged_files_rejected(tmp_path, rich):
    path = str(tmp_path / "g.lvlz")
    capi.levelize_to_file(rich, path)
    good = bytearray(open(path, "rb").read())
    bad = str(tmp_path / "bad.lvlz")

    def expect(buf, match):
        open(bad, "wb").write(bytes(buf))
        with pytest.raises(capi.GrglError, match=match):
            capi.levelized_file_info(bad)

    flipped = bytearray(good)
    flipped[len(good) // 2] ^= 0x40
    expect(flipped, "checksum")
    expect(good[:-9], "truncated|trailing|checksum")
    expect(good[: len(good) // 3], "truncated")
    expect(good + b"\0" * 8, "trailing")
    wrong = bytearray(good)
    wrong[0:8] = b"GRGLXXXX"
    expect(wrong, "magic")
    newer = bytearray(good)
    newer[8:12] = struct.pack("<I", 99)
    expect(newer, "version")
    with pytest.raises(capi.GrglError, match="cannot open"):
        capi.levelized_file_info(str(tmp_path / "missing.lvlz"))


def test_invalid_plans_rejected_despite_checksum(tmp_path, rich):
    """Files whose checksum is right but whose plans break what the kernels rely on."""
    path = str(tmp_path / "g.lvlz")
    capi.levelize_to_file(rich, path)
    good = bytearray(open(path, "rb").read())
    up, down = _parse(good)
    V = len(rich["down_offsets"]) - 1
    M = len(rich["mutation_nodes"])
    bad = str(tmp_path / "bad.lvlz")

    def expect(mutate, match):
        buf = bytearray(good)
        mutate(buf)
        _rewrite(bad, buf)
        with pytest.raises(capi.GrglError, match=match):
            capi.levelized_file_info(bad)

    def put(buf, field, index, value):
        struct.pack_into("<I", buf, field[0] + 4 * index, value)

    # an UP entry naming a row of its own level (read before it is written)
    ls = _array(good, up["levelStart"])
    off = _array(good, up["off"])
    top = len(ls) - 2
    j = int(ls[top])
    assert off[j + 1] > off[j]
    expect(lambda b: put(b, up["src"], int(off[j]), j), "earlier levels")
    # a DOWN entry past the seed rows (V + 2M with missing data)
    expect(lambda b: put(b, down["src"], 0, V + 2 * M), "earlier levels")
    # rows of a level not longest list first
    def unsort(b):
        o = _array(good, up["off"])
        lvl_rows = [(l, int(ls[l]), int(ls[l + 1])) for l in range(1, len(ls) - 1)]
        for _, a, e in lvl_rows:
            for r in range(a, e - 1):
                if o[r + 1] - o[r] > o[r + 2] - o[r + 1]:
                    # make row r shorter by one and row r+1 longer by one: still covering the
                    # same entries, now increasing
                    put(b, up["off"], r + 1, int(o[r + 1]) - (o[r + 1] - o[r]) + (o[r + 2] - o[r + 1]))
                    return
        raise AssertionError("no level with two different list lengths")
    expect(unsort, "longest first|earlier levels")
    # row order not a permutation
    order = _array(good, up["order"])
    expect(lambda b: put(b, up["order"], V - 1, int(order[V - 2])), "permutation")
    # maximum fan-in field
    expect(lambda b: struct.pack_into("<I", b, up["scalars"] + 12, 1), "maximum fan-in")
    # read-out rows out of range (UP: mutation rows, DOWN: sample rows)
    expect(lambda b: put(b, up["outRow"], 0, V + 5), "read-out")
    expect(lambda b: put(b, down["outRow"], 0, 0x7FFFFFFF), "read-out")
    expect(lambda b: put(b, up["missRow"], 1, V), "missingness")
    # level bounds
    expect(lambda b: put(b, up["levelStart"], len(ls) - 1, V - 1), "level bounds")
    # a sample's coalescence count
    expect(lambda b: put(b, up["coalByRow"], 0, 3), "coalescence")
