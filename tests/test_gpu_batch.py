"""GPU: uneven batches in persistent launches (CkFftBatchPlan*) and the autotuner.

Every transform of a plan must equal, bit for bit, the same transform run alone through
the single-signal drop-in call (same arithmetic by construction), and match the CPU
oracle within the parity gates of test_gpu_parity.py.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

KINDS = ("cfwd", "cinv", "rfwd", "rinv")


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


@pytest.fixture(scope="module")
def C(torch):
    import ckfft_amd
    return ckfft_amd


@pytest.fixture(scope="module")
def ctx(torch, C):
    c = C.Context(1 << 17, C.kCkFftDirection_Both)
    yield c
    c.close()


def in_len(kind, n):
    return {"cfwd": 2 * n, "cinv": 2 * n, "rfwd": n, "rinv": 2 * (n // 2 + 1)}[kind]


def out_len(kind, n):
    return {"cfwd": 2 * n, "cinv": 2 * n, "rfwd": 2 * (n // 2 + 1), "rinv": n}[kind]


def make_input(kind, n, seed):
    if kind in ("cfwd", "cinv"):
        return O.uniform_pm1(seed, 2 * n)
    if kind == "rfwd":
        return O.normal01(seed, n)
    return O.Restatement(n, 3).real_forward(O.normal01(seed, n))


def plain(torch, ctx, kind, n, xt):
    y = torch.full((out_len(kind, n),), float("nan"), dtype=torch.float32, device="cuda")
    if kind == "cfwd":
        ok = ctx.complex_forward(n, xt, y)
    elif kind == "cinv":
        ok = ctx.complex_inverse(n, xt, y)
    elif kind == "rfwd":
        ok = ctx.real_forward(n, xt, y)
    else:
        ok = ctx.real_inverse(n, xt, y, torch.empty(2 * (n // 2 + 1), device="cuda"))
    assert ok == 1
    return y


def random_mix(count, seed, kmin=0, kmax=17):
    rng = np.random.default_rng(seed)
    return [(KINDS[rng.integers(0, 4)], 1 << int(rng.integers(kmin, kmax + 1))) for _ in range(count)]


def build(torch, C, ctx, mix, seed0):
    ins, outs, descs = [], [], []
    for i, (kind, n) in enumerate(mix):
        x = torch.from_numpy(np.ascontiguousarray(make_input(kind, n, seed0 + i), dtype=np.float32)).cuda()
        y = torch.full((out_len(kind, n),), float("nan"), dtype=torch.float32, device="cuda")
        ins.append(x)
        outs.append(y)
        descs.append((KINDS.index(kind), n, x, y))
    return ins, outs, descs


def test_uneven_batch_equals_single_calls(torch, C, ctx):
    mix = random_mix(160, 11)
    mix += [("cfwd", 1 << 16), ("cinv", 1 << 16), ("rfwd", 1 << 17), ("cfwd", 1 << 17), ("rinv", 4), ("cfwd", 1)]
    ins, outs, descs = build(torch, C, ctx, mix, 1000)
    plan = C.BatchPlan(ctx, descs)
    info = plan.info()
    assert info["transforms"] == len(mix)
    assert info["extra_launches"] >= 3          # the closed-form sizes and 2^17
    # complex 2^16: table-driven pass launches
    assert info["pass_transforms"] == sum(k in ("cfwd", "cinv") and n == 1 << 16 for k, n in mix)
    assert plan.execute() == 1
    torch.cuda.synchronize()
    for i, (kind, n) in enumerate(mix):
        want = plain(torch, ctx, kind, n, ins[i]).cpu().numpy()
        got = outs[i].cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (i, kind, n)
    plan.close()


def test_uneven_batch_against_oracle(torch, C, ctx):
    mix = random_mix(60, 12, kmin=3, kmax=16)
    ins, outs, descs = build(torch, C, ctx, mix, 2000)
    plan = C.BatchPlan(ctx, descs)
    assert plan.execute() == 1
    torch.cuda.synchronize()
    for i, (kind, n) in enumerate(mix):
        x = ins[i].cpu().numpy()[None]
        o = O.Restatement(n, 3)
        ref = {"cfwd": lambda: o.complex(x, False), "cinv": lambda: o.complex(x, True),
               "rfwd": lambda: o.real_forward(x), "rinv": lambda: o.real_inverse(x, n)}[kind]()
        got = outs[i].cpu().numpy()[None]
        assert O.split_norm_err(got, ref) <= 1e-5, (i, kind, n)


def test_graph_replays_follow_inputs(torch, C, ctx):
    """A graph-captured plan replays correctly with eager work between replays, and each
    replay reads the inputs as they are at replay time."""
    mix = random_mix(40, 13, kmin=4, kmax=16)
    ins, outs, descs = build(torch, C, ctx, mix, 3000)
    plan = C.BatchPlan(ctx, descs, graph=True)
    for rep in range(3):
        for i, (kind, n) in enumerate(mix):
            if kind in ("cfwd", "cinv", "rfwd"):
                ins[i].mul_(1.5)                 # eager work that changes the inputs
        assert plan.execute() == 1
        torch.cuda.synchronize()
        for i, (kind, n) in enumerate(mix):
            want = plain(torch, ctx, kind, n, ins[i]).cpu().numpy()
            assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (rep, i, kind, n)


def test_autotune_changes_speed_not_results(torch, C, tmp_path):
    with C.Context(1 << 16) as c:
        path = str(tmp_path / "tune.json")
        assert c.autotune(path) == 1
        data = json.load(open(path))
        # every class with more than one layout was timed: complex 2^5..2^14 and 2^16, real
        # 2^6..2^15 samples; each choice is a valid code
        assert data["arch"] == "gfx950" and data["format"] == 2 and len(data["us"]) == 2 * 11 + 2 * 10
        assert all(1 <= v <= 4 for v in data["choice"].values())
        mix = random_mix(80, 14, kmin=5, kmax=14)
        ins, outs, descs = build(torch, C, c, mix, 4000)
        plan = C.BatchPlan(c, descs)
        assert plan.execute() == 1
        torch.cuda.synchronize()
        for i, (kind, n) in enumerate(mix):
            want = plain(torch, c, kind, n, ins[i]).cpu().numpy()
            assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)
        with C.Context(1 << 16) as c2:
            assert c2.load_tuning(path) == 1


def test_wide_and_narrow_queues_equal_single_calls(torch, C, tmp_path):
    """Every class through both persistent kernels: with the 32-points-per-thread layout
    chosen for every tunable class (a tuning file), the classes up to 2^13 complex points
    run in the wide kernel and 2^14 in the narrow one; with the default layout they run in
    the narrow one.  Either way each output equals the single call bit for bit."""
    path = tmp_path / "all_var1.json"
    keys = [f'"{k}_{(c + 1) if k[0] == "r" else c}": 1' for k in KINDS for c in range(5, 14)]
    path.write_text('{\n "arch": "gfx950",\n "choice": {' + ", ".join(keys) + "}\n}\n")
    mix = [(k, 1 << e) for k in KINDS for e in range(1 if k[0] == "c" else 3, 16 if k[0] == "c" else 17)]
    for tuned in (False, True):
        with C.Context(1 << 16) as c:
            if tuned:
                assert c.load_tuning(str(path)) == 1
            ins, outs, descs = build(torch, C, c, mix + mix, 5000)
            plan = C.BatchPlan(c, descs)
            assert plan.info()["persistent_transforms"] == len(descs)
            assert plan.execute() == 1
            torch.cuda.synchronize()
            for i, (kind, n) in enumerate(mix + mix):
                want = plain(torch, c, kind, n, ins[i]).cpu().numpy()
                assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (tuned, i, kind, n)
            plan.close()


def test_queue_routing_by_class(torch, C):
    """Classes up to 2^13 complex points at 16 points per thread, and 2^14, go to the narrow
    kernel (two workgroups per CU); 2^15 complex and 2^16 real go to the wide one (one per
    CU).  The grid sizes CkFftBatchInfo reports follow the items of each queue."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    # (transforms, wide grid, narrow grid); a grid is min(items, CUs x workgroups per CU), an
    # item holds 512 / (n / points per thread) transforms of one class
    cases = [
        ([("cfwd", 1 << 12)] * 3 * cus, 0, min(3 * cus // 2, 2 * cus)),  # 2 per item, narrow
        ([("rinv", 1 << 15)] * 3 * cus, 0, 2 * cus),                     # real 2^15: 2^14 points, narrow
        ([("cinv", 1 << 15)] * 3 * cus, cus, 0),                         # wide only
        ([("rfwd", 1 << 16)] * 5, 5, 0),
        ([("cfwd", 1 << 15)] * 3 + [("cfwd", 1 << 10)] * 8, 3, 1),       # 8 transforms of 2^10: one item
    ]
    with C.Context(1 << 16) as c:
        for mix, wide, narrow in cases:
            ins, outs, descs = build(torch, C, c, mix, 9100)
            plan = C.BatchPlan(c, descs)
            info = plan.info()
            assert info["workgroups"] == wide + narrow, (mix[0], info)
            assert plan.execute() == 1
            torch.cuda.synchronize()
            for i in (0, len(mix) - 1):
                kind, n = mix[i]
                want = plain(torch, c, kind, n, ins[i]).cpu().numpy()
                assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)
            plan.close()


def test_batch_plan_validation(torch, C, ctx):
    x = torch.zeros(64, device="cuda")
    y = torch.zeros(64, device="cuda")
    h = np.zeros(64, np.float32)
    bad = [
        [(0, 16, x, x)],                 # in == out
        [(0, 12, x, y)],                 # not a power of two
        [(0, 1 << 18, x, y)],            # above nMax
        [(5, 16, x, y)],                 # bad kind
        [(0, 16, h, y)],                 # host input
    ]
    for descs in bad:
        with pytest.raises(ValueError):
            C.BatchPlan(ctx, descs)
    with C.Context(64, C.kCkFftDirection_Forward) as f:
        with pytest.raises(ValueError):
            C.BatchPlan(f, [(1, 16, x, y)])  # inverse on a forward-only context
    p = C.BatchPlan(ctx, [])
    assert p.execute() == 1


def test_config_uneven_sample_2000(torch, C, ctx):
    """A 2000-transform sample of the uneven workload (k in 6..16, kinds and directions
    50/50, N(0,1) data): every output against the single calls, bitwise."""
    rng = np.random.default_rng(0xC4)
    mix = []
    for _ in range(2000):
        k = int(rng.integers(6, 17))
        real = rng.random() < 0.5
        fwd = rng.random() < 0.5
        kind = ("rfwd" if fwd else "rinv") if real else ("cfwd" if fwd else "cinv")
        mix.append((kind, 1 << k))
    ins, outs, descs = build(torch, C, ctx, mix, 5000)
    plan = C.BatchPlan(ctx, descs, graph=True)
    info = plan.info()
    assert info["extra_launches"] == 0 and info["pass_transforms"] > 0
    assert info["persistent_transforms"] + info["pass_transforms"] == len(mix)
    assert plan.execute() == 1
    torch.cuda.synchronize()
    for i in range(0, 2000, 7):
        kind, n = mix[i]
        want = plain(torch, ctx, kind, n, ins[i]).cpu().numpy()
        assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)


def test_config_uneven_full_20000(torch, C, ctx):
    """The whole uneven workload: 20,000 transforms, n = 2^k with k uniform in 6..16,
    kinds and directions 50/50, N(0,1) data, each with its own 256-byte aligned buffers,
    in one graph-replayed plan (run twice).  A bitwise sample of every 50th transform
    against the single calls, and every 1,000th against the CPU oracle."""
    count = 20000
    rng = np.random.default_rng(0xC4)
    ks = rng.integers(6, 17, size=count)
    real = rng.random(count) < 0.5
    fwd = rng.random(count) < 0.5
    kinds = np.where(real, np.where(fwd, 2, 3), np.where(fwd, 0, 1))
    ns = (1 << ks).astype(np.int64)
    il = np.where(kinds == 2, ns, np.where(kinds == 3, 2 * (ns // 2 + 1), 2 * ns))
    ol = np.where(kinds == 2, 2 * (ns // 2 + 1), np.where(kinds == 3, ns, 2 * ns))
    ioff = np.concatenate([[0], np.cumsum((il + 63) // 64 * 64)])
    ooff = np.concatenate([[0], np.cumsum((ol + 63) // 64 * 64)])
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0xC4)
    x = torch.randn(int(ioff[-1]), device="cuda", generator=gen)
    y = torch.full((int(ooff[-1]),), float("nan"), device="cuda")
    xs = [x[int(ioff[i]): int(ioff[i]) + int(il[i])] for i in range(count)]
    ys = [y[int(ooff[i]): int(ooff[i]) + int(ol[i])] for i in range(count)]
    descs = [(int(kinds[i]), int(ns[i]), xs[i], ys[i]) for i in range(count)]
    print("uneven 20000: building plan", flush=True)
    plan = C.BatchPlan(ctx, descs, graph=True)
    info = plan.info()
    assert info["transforms"] == count and info["extra_launches"] == 0
    assert info["persistent_transforms"] + info["pass_transforms"] == count
    for _ in range(2):
        assert plan.execute() == 1
    torch.cuda.synchronize()
    # every output value written: the only NaNs left are the alignment gaps between rows
    assert int(torch.isnan(y).sum()) == int(ooff[-1]) - int(ol.sum())
    for i in range(0, count, 50):
        want = plain(torch, ctx, KINDS[kinds[i]], int(ns[i]), xs[i]).cpu().numpy()
        assert np.array_equal(ys[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kinds[i], ns[i])
    for i in range(0, count, 1000):
        kind, n = KINDS[kinds[i]], int(ns[i])
        xi = xs[i].cpu().numpy()[None]
        o = O.Restatement(n, 3)
        ref = (o.complex(xi, kind == "cinv") if kind in ("cfwd", "cinv")
               else o.real_forward(xi) if kind == "rfwd" else o.real_inverse(xi, n))
        assert O.split_norm_err(ys[i].cpu().numpy()[None], ref) <= 1e-5, (i, kind, n)
    plan.close()


def test_every_layout_code_equals_single_calls(torch, C, tmp_path):
    """A tuning file (second format) with a non-default layout for every class: the small
    classes in the narrow kernel, 2^11..2^14 in the wide one (2^14 then with both LDS planes
    at once), and complex 2^16 as ordinary two-pass launches instead of the table passes.
    Every output still equals the single call bit for bit."""
    keys = []
    for k in KINDS:
        for c in range(1 if k[0] == "c" else 2, 15):
            code = 1 if c <= 10 else 3               # var 0: narrow for the small, wide for the rest
            keys.append(f'"{k}_{(c + 1) if k[0] == "r" else c}": {code}')
    keys += ['"cfwd_16": 2', '"cinv_16": 2']
    path = tmp_path / "every_code.json"
    path.write_text('{\n "arch": "gfx950",\n "format": 2,\n "choice": {' + ", ".join(keys) + "}\n}\n")
    mix = [(k, 1 << e) for k in KINDS for e in range(1 if k[0] == "c" else 3, 17 if k[0] == "c" else 17)]
    with C.Context(1 << 16) as c:
        assert c.load_tuning(str(path)) == 1
        ins, outs, descs = build(torch, C, c, mix + mix, 7000)
        plan = C.BatchPlan(c, descs)
        info = plan.info()
        assert info["pass_transforms"] == 0                  # complex 2^16 as ordinary launches
        assert info["extra_launches"] == sum(1 for k, n in mix + mix if (k[0] == "c" and n in (1, 1 << 16))
                                             or (k[0] == "r" and n <= 4))
        assert plan.execute() == 1
        torch.cuda.synchronize()
        for i, (kind, n) in enumerate(mix + mix):
            want = plain(torch, c, kind, n, ins[i]).cpu().numpy()
            assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)
        plan.close()


def test_autotune_mix_keeps_results(torch, C, tmp_path):
    """CkFftAutotuneMix over an uneven mix times the whole plan with each class's other
    layouts; it never leaves layouts slower than those on entry, and the tuned plan's
    outputs equal the single calls bit for bit."""
    rng = np.random.default_rng(0x7E)
    mix = []
    for _ in range(600):
        k = int(rng.integers(5, 17))
        real, fwd = rng.random() < 0.5, rng.random() < 0.5
        mix.append((("rfwd" if fwd else "rinv") if real else ("cfwd" if fwd else "cinv"), 1 << k))
    with C.Context(1 << 16) as c:
        ins, outs, descs = build(torch, C, c, mix, 7500)
        path = str(tmp_path / "mix.json")
        ok, before, after = c.autotune_mix(descs, path)
        assert ok == 1 and before > 0 and 0 < after <= before
        data = json.load(open(path))
        assert data["format"] == 2 and len(data["us"]) > 0
        plan = C.BatchPlan(c, descs, graph=True)
        assert plan.execute() == 1
        torch.cuda.synchronize()
        for i, (kind, n) in enumerate(mix):
            want = plain(torch, c, kind, n, ins[i]).cpu().numpy()
            assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)
        plan.close()


def test_plan_branches_join_the_callers_stream(torch, C, ctx):
    """A plan runs its table passes, its two-per-CU persistent launch and its one-per-CU
    persistent launch side by side, on streams of its own forked from the caller's stream.
    Work queued on the caller's stream after an execution must see every output of every
    branch: the outputs are cleared and snapshotted on that stream around each execution,
    eagerly and as graph replays, and every snapshot must equal the single calls bit for
    bit."""
    mix = ([("cfwd", 1 << 16)] * 3 + [("cinv", 1 << 16)] * 2 + [("rfwd", 1 << 16)] * 6 +
           [("rinv", 1 << 16)] * 4 + [("cinv", 4096)] * 24 + [("rfwd", 2048)] * 24 + [("cfwd", 256)] * 8)
    ins, outs, descs = build(torch, C, ctx, mix, 9100)
    want = [plain(torch, ctx, kind, n, ins[i]).cpu().numpy() for i, (kind, n) in enumerate(mix)]
    s = torch.cuda.Stream()
    for graph in (False, True):
        plan = C.BatchPlan(ctx, descs, graph=graph)
        info = plan.info()
        assert info["pass_transforms"] == 5 and info["persistent_transforms"] == len(mix) - 5
        for rep in range(3):
            with torch.cuda.stream(s):
                for y in outs:
                    y.fill_(float("nan"))
            assert plan.execute(s) == 1
            with torch.cuda.stream(s):
                snap = [y.clone() for y in outs]
            torch.cuda.synchronize()
            for i, (kind, n) in enumerate(mix):
                got = snap[i].cpu().numpy()
                assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (graph, rep, i, kind, n)
        plan.close()
