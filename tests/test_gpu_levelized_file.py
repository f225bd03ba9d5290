"""Levelized-graph files on the device (grgl_hip_graph_store / grgl_hip_graph_load, the
reference's storeGPUGRGToDisk / loadGPUGRGFromDisk, src/cuda/gpu_grg.cu:10-141): a graph
stored and loaded again, or levelized on the host and loaded, multiplies bit for bit like the
graph created from the arrays, for integer and float inputs, both directions, with
missingness, XTX initialisation and batches; a damaged file is refused before any device
allocation."""

import numpy as np
import pytest

import grgl_amd
import oracle
from grgl_amd import capi
from helpers import bits_equal, make_input, random_grg

pytestmark = pytest.mark.gpu


def _graphs():
    src = grgl_amd.grg_from_coalescent(1000, 10000, 2, seed=3)
    off, car = src.mutation_carriers()
    return {
        "rich": random_grg(9, num_samples=60, num_internal=300, max_children=7, heavy=(80, 200), num_mutations=400,
                           with_miss=True, with_coal=True),
        "coalescent": src.arrays(),
        "constructed": grgl_amd.fast_grg_from_samples(1000, 2, carrier_offsets=off, carriers=car).arrays(),
    }


def _run(cg, dt, x, direction, **kw):
    import torch

    K, cin = x.shape
    cout = cg.num_mutations if direction == capi.UP else cg.num_samples
    xd = torch.from_numpy(x).cuda()
    y = torch.zeros((K, cout), dtype=xd.dtype, device="cuda")
    cg.matmul_ptr(dt, xd.data_ptr(), cin, K, direction, y.data_ptr(), y.numel(),
                  stream=torch.cuda.current_stream().cuda_stream, **kw)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("how", ["store", "host"])
def test_loaded_graph_multiplies_bit_for_bit(gpu, tmp_path, how):
    import torch

    rng = np.random.default_rng(17)
    for name, arrays in _graphs().items():
        made = capi.Graph(arrays)
        path = str(tmp_path / f"{name}.lvlz")
        if how == "store":
            made.store(path)
        else:
            capi.levelize_to_file(arrays, path)
        loaded = capi.Graph.load(path)
        a, b = made.info(), loaded.info()
        for k in ("num_samples", "num_nodes", "num_mutations", "ploidy", "num_edges", "up_levels", "down_levels",
                  "max_fanin_up", "max_fanin_down", "up_gather_entries", "down_gather_entries", "device_bytes"):
            assert a[k] == b[k], (name, k)
        assert capi.levelized_file_info(path)["device_bytes"] == a["device_bytes"]
        og = oracle.OracleGraph(arrays)
        for direction in (capi.UP, capi.DOWN):
            cin = made.num_samples if direction == capi.UP else made.num_mutations
            for dtype, K in ((np.int32, 1), (np.int32, 16), (np.float32, 5), (np.float64, 3)):
                dt = capi.NUMPY_TO_DTYPE[np.dtype(dtype)]
                x = make_input(rng, dtype, (K, cin))
                want = _run(made, dt, x, direction)
                assert bits_equal(_run(loaded, dt, x, direction), want), (name, direction, dtype, K)
                if dtype == np.int32:
                    assert bits_equal(want, og.matmul(x, oracle.UP if direction == capi.UP else oracle.DOWN))
            if name == "rich" and direction == capi.UP:
                # XTX initialisation (coalescence counts) and the missingness output
                x = make_input(rng, np.int64, (2, cin))
                dt = capi.I64
                assert bits_equal(_run(loaded, dt, x, direction, init_mode=capi.INIT_XTX),
                                  _run(made, dt, x, direction, init_mode=capi.INIT_XTX))
                outs = []
                for cg in (made, loaded):
                    miss = torch.zeros((2, cg.num_mutations), dtype=torch.int64, device="cuda")
                    _run(cg, dt, x, direction, miss_ptr=miss.data_ptr())
                    outs.append(miss.cpu().numpy())
                assert bits_equal(outs[0], outs[1])
        # a batch on the loaded graph
        xs = [torch.from_numpy(make_input(rng, np.int32, (k, made.num_samples))).cuda() for k in (1, 3, 4)]
        res = []
        for cg in (made, loaded):
            outs = [torch.zeros((x.shape[0], cg.num_mutations), dtype=torch.int32, device="cuda") for x in xs]
            cg.matmul_batch_ptr(capi.I32, capi.UP, [(x.shape[0], x.data_ptr(), o.data_ptr()) for x, o in zip(xs, outs)],
                                stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            res.append([o.cpu().numpy() for o in outs])
        for p, q in zip(*res):
            assert bits_equal(p, q)
        made.close()
        loaded.close()


def test_damaged_file_refused(gpu, tmp_path):
    arrays = _graphs()["rich"]
    path = str(tmp_path / "g.lvlz")
    capi.levelize_to_file(arrays, path)
    buf = bytearray(open(path, "rb").read())
    buf[len(buf) // 2] ^= 1
    open(path, "wb").write(bytes(buf))
    with pytest.raises(capi.GrglError, match="checksum"):
        capi.Graph.load(path)
    with pytest.raises(capi.GrglError, match="cannot open"):
        capi.Graph.load(str(tmp_path / "none.lvlz"))
