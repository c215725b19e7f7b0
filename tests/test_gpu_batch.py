"""Batched form on the GPU: one persistent launch over many uneven items gives,
for every item, exactly the stream of a plain call on that item (and the
oracle's), also when replayed from a hipGraph and when the plan is refilled."""
import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu


def _frames(seed, n, sides=(8, 24, 64, 136, 256, 512)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        y = int(rng.choice(sides)) + int(rng.integers(0, 3))
        z = int(rng.choice(sides)) + int(rng.integers(0, 2)) * 2
        x = int(rng.integers(1, 3))
        out.append(synth("xpcs" if i % 2 else "cssi", (x, y, z), seed=seed * 1000 + i))
    return out


def _compress_batch(gpu, vols, params, graph=False, plan=None):
    import torch
    dev = [torch.from_numpy(v.view(np.int16)).cuda() for v in vols]
    outs = [torch.zeros(gpu.max_compressed_bytes(v.shape), dtype=torch.uint8, device="cuda") for v in vols]
    if plan is None:
        plan = gpu.Plan(gpu.COMPRESS, len(vols), sum(gpu.num_blocks(v.shape) for v in vols))
    plan.set_items([(d, v.shape, b, p, o, o.numel()) for d, v, (b, p), o in zip(dev, vols, params, outs)])
    plan.launch(graph=graph)
    sizes = plan.sizes()
    assert plan.status() == gpu.OK
    return plan, [o[:n].cpu().numpy() for o, n in zip(outs, sizes)]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle_and_plain(gpu, orc, graph):
    vols = _frames(1, 24)
    params = [((1, 1, 1, 1), 1.0) if i % 2 == 0 else ((3, 5, 10, 15), 0.5) for i in range(len(vols))]
    plan, streams = _compress_batch(gpu, vols, params, graph=graph)
    for v, (b, p), s in zip(vols, params, streams):
        st, ref, refn = orc.compress(v, b, p)
        assert s.size == refn and np.array_equal(s, ref)
    # decompress the same streams as one batch
    import torch
    dplan = gpu.Plan(gpu.DECOMPRESS, len(vols), sum(gpu.num_blocks(v.shape) for v in vols))
    dsrc = [torch.from_numpy(s.copy()).cuda() for s in streams]
    dout = [torch.full(v.shape, 7, dtype=torch.int16, device="cuda") for v in vols]
    dplan.set_items([(o, v.shape, b, p, s, s.numel()) for o, v, (b, p), s in zip(dout, vols, params, dsrc)])
    dplan.launch(graph=graph)
    assert dplan.status() == gpu.OK
    for o, v, (b, p), s in zip(dout, vols, params, streams):
        _, refdec = orc.decompress(s, v.shape, b)
        assert np.array_equal(o.cpu().numpy().view(np.uint16), refdec)
    plan.close()
    dplan.close()


def test_plan_refill_and_graph_replay(gpu, orc):
    """One plan, several batches; the captured graph picks up each new batch."""
    vols_all = _frames(2, 30)
    maxb = sum(gpu.num_blocks(v.shape) for v in vols_all)
    plan = gpu.Plan(gpu.COMPRESS, 30, maxb)
    for rep, (lo, hi) in enumerate([(0, 10), (10, 30), (5, 12), (0, 30)]):
        vols = vols_all[lo:hi]
        params = [((3, 5, 10, 15), 0.5)] * len(vols)
        _, streams = _compress_batch(gpu, vols, params, graph=True, plan=plan)
        for v, s in zip(vols, streams):
            st, ref, refn = orc.compress(v, (3, 5, 10, 15), 0.5)
            assert np.array_equal(s, ref), (rep, v.shape)
    plan.close()


def test_decode_plan_refill_and_graph_replay(gpu, orc):
    """One decompression plan over several batches (graph replays, items of
    1 to 75 tiles, a corrupt stream in one batch): the offset phase runs over
    every item's tiles and each item decodes to the oracle's volume; the
    corrupt item is reported and the next batch decodes cleanly."""
    import torch
    vols_all = _frames(3, 30, sides=(8, 64, 136, 512, 1100))
    params_all = [((1, 1, 1, 1), 1.0) if i % 3 else ((3, 5, 10, 15), 0.5) for i in range(len(vols_all))]
    refs = [orc.compress(v, b, p)[1] for v, (b, p) in zip(vols_all, params_all)]
    maxb = sum(gpu.num_blocks(v.shape) for v in vols_all)
    dplan = gpu.Plan(gpu.DECOMPRESS, 30, maxb)
    for rep, (lo, hi) in enumerate([(0, 10), (10, 30), (5, 12), (0, 30), (3, 4)]):
        vols, params, streams = vols_all[lo:hi], params_all[lo:hi], refs[lo:hi]
        src = [torch.from_numpy(r.copy()).cuda() for r in streams]
        outs = [torch.full(v.shape, 7, dtype=torch.int16, device="cuda") for v in vols]
        dplan.set_items([(o, v.shape, b, p, sb, sb.numel()) for o, v, (b, p), sb in zip(outs, vols, params, src)])
        dplan.launch(graph=True)
        assert dplan.status() == gpu.OK, rep
        for o, v, (b, p), r in zip(outs, vols, params, streams):
            _, refdec = orc.decompress(r, v.shape, b)
            assert np.array_equal(o.cpu().numpy().view(np.uint16), refdec), (rep, v.shape)
    # a header byte with rate 31 in one item: flagged, that item's bad block decodes to zeros
    bad = refs[4].copy()
    bad[0] = 0x1f
    vols, params = vols_all[:6], params_all[:6]
    src = [torch.from_numpy((bad if i == 4 else refs[i]).copy()).cuda() for i in range(6)]
    outs = [torch.full(v.shape, 7, dtype=torch.int16, device="cuda") for v in vols]
    dplan.set_items([(o, v.shape, b, p, sb, sb.numel()) for o, v, (b, p), sb in zip(outs, vols, params, src)])
    dplan.launch(graph=True)
    assert dplan.status() == gpu.ERR_CORRUPT
    for i in (0, 1, 2, 3, 5):
        _, refdec = orc.decompress(refs[i], vols[i].shape, params[i][0])
        assert np.array_equal(outs[i].cpu().numpy().view(np.uint16), refdec)
    src[4] = torch.from_numpy(refs[4].copy()).cuda()
    dplan.set_items([(o, v.shape, b, p, sb, sb.numel()) for o, v, (b, p), sb in zip(outs, vols, params, src)])
    dplan.launch(graph=True)
    assert dplan.status() == gpu.OK
    _, refdec = orc.decompress(refs[4], vols[4].shape, params[4][0])
    assert np.array_equal(outs[4].cpu().numpy().view(np.uint16), refdec)
    dplan.close()


def test_plan_rejects_bad_items(gpu):
    import torch
    plan = gpu.Plan(gpu.COMPRESS, 2, 64)
    v = torch.zeros((1, 8, 8), dtype=torch.int16, device="cuda")
    o = torch.zeros(256, dtype=torch.uint8, device="cuda")
    with pytest.raises(gpu.LscompError) as ei:
        plan.set_items([(v, (1, 8, 8), (3, 2, 5, 6), 0.5, o, 256)])
    assert ei.value.status == gpu.ERR_INVALID_ARG
    big = torch.zeros((1, 1024, 1024), dtype=torch.int16, device="cuda")
    with pytest.raises(gpu.LscompError) as ei:      # more blocks than the plan holds
        plan.set_items([(big, (1, 1024, 1024), (1, 1, 1, 1), 1.0, o, 256)])
    plan.close()


def test_plan_shape_change_and_autotune(gpu, orc):
    """A plan's shape can change after its batch is set (tile tables rebuilt,
    captured graph recaptured), and the plan autotuner, timing the plan's own
    batch, applies one of the shapes it measured; every stream stays the
    oracle's."""
    import torch
    vols = _frames(5, 20)
    params = [((1, 1, 1, 1), 1.0) if i % 2 == 0 else ((3, 5, 10, 15), 0.5) for i in range(len(vols))]
    dev = [torch.from_numpy(v.view(np.int16)).cuda() for v in vols]
    outs = [torch.zeros(gpu.max_compressed_bytes(v.shape), dtype=torch.uint8, device="cuda") for v in vols]
    refs = [orc.compress(v, b, p)[1] for v, (b, p) in zip(vols, params)]
    plan = gpu.Plan(gpu.COMPRESS, len(vols), sum(gpu.num_blocks(v.shape) for v in vols))
    plan.set_items([(d, v.shape, b, p, o, o.numel()) for d, v, (b, p), o in zip(dev, vols, params, outs)])

    def check():
        sizes = plan.sizes()
        assert plan.status() == gpu.OK
        for o, n, ref in zip(outs, sizes, refs):
            assert n == ref.size and np.array_equal(o[:n].cpu().numpy(), ref)
            o.zero_()

    plan.launch(graph=True)
    check()
    for shape in ((128, 1, 128, 2), (256, 0, 64, 16), (512, 0, 256, 1)):
        plan.set_shape(*shape)
        assert plan.get_shape() == dict(zip(("tile", "wg_per_cu", "lb_window", "counters"), shape))
        plan.launch(graph=True)
        check()
    table = plan.autotune(reps=1)
    assert {e["tile"] for e in table} == {128, 256, 512} and all(e["direction"] == 0 for e in table)
    cur = plan.get_shape()
    assert any(all(e[f] == cur[f] for f in cur) for e in table)
    for o in outs:
        o.zero_()
    plan.launch(graph=True)
    check()
    plan.close()
