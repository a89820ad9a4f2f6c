"""Batched eigen solver (pk_batch_*): every member of a mixed batch (sizes,
group counts, boundary conditions) against the FP64 oracle within |dk| <= 1 pcm
and scalar-flux relative L2 <= 1e-5, in FP64 and FP32-inner precision; and each
member's result bit-identical whether it is solved alone or in the batch."""
import os

import numpy as np
import pytest

from conftest import REF

pytestmark = pytest.mark.gpu

DK_PCM = 1.0
FLUX_L2 = 1e-5

# (name, batched-instance index or reference case)
MEMBERS = [("slab-1d", None), ("inst-0", 0), ("inst-3", 3), ("inst-5", 5), ("inst-9", 9),
           ("inst-20", 20), ("inst-28", 28), ("reflected", "slabs/reflected-diffusion"),
           ("pwr-cart", "pwr-iaea-benchmark/cartesian-diffusion")]


@pytest.fixture(scope="module")
def members(oracle_mod, tmp_path_factory):
    from pampa_amd import workloads as W
    base = tmp_path_factory.mktemp("batch")
    out = []
    for name, src in MEMBERS:
        d = str(base / name)
        if name == "slab-1d":
            W.slab_1d(d)
        elif isinstance(src, int):
            W.batched_instance(d, src)
        else:
            d = os.path.join(REF, src)
        cwd = os.getcwd()
        os.chdir(d)
        try:
            case = oracle_mod.pmp.read_input("input.pmp")
            ref = oracle_mod.solve_input("input.pmp", return_ops=True)
        finally:
            os.chdir(cwd)
        out.append((name, case.solver.G, ref))
    return out


_MIX = {}  # precision -> [(keff, eigenvector)] of the fixed batch


def _mats(pk, members):
    Rs, Fs = [], []
    for _, G, ref in members:
        Rs.append(pk.Mat.from_csr(ref.R.n, ref.R.ptr, ref.R.col, ref.R.val, block_size=G))
        Fs.append(pk.Mat.from_csr(ref.F.n, ref.F.ptr, ref.F.col, ref.F.val, block_size=G))
    return Rs, Fs


def _shape(v):
    v = np.asarray(v, dtype=np.float64)
    return v / v.sum()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_parity_and_alone_vs_mix(pk, members, precision):
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    Rs, Fs = _mats(pk, members)
    print(f"[case] batch of {len(Rs)} {precision}", flush=True)
    batch = pk.Batch(Rs, Fs, precision=prec)
    batch.solve()
    mix = []
    for b, (name, G, ref) in enumerate(members):
        k, x = batch.eigenpair(b)
        st = batch.stats(b)
        xv = x.get()
        dk = abs(k - ref.keff) * 1e5
        l2 = np.linalg.norm(_shape(xv) - _shape(ref.phi)) / np.linalg.norm(_shape(ref.phi))
        print(f"   {name}: G={G} n={ref.R.n} dk={dk:.2e} pcm l2={l2:.2e} outer={st['outer']} "
              f"inner={st['inner']}", flush=True)
        assert dk <= DK_PCM, name
        assert l2 <= FLUX_L2, name
        mix.append((k, xv))
    _MIX[precision] = mix
    # a second solve of the same batch (graph replays) gives the same bits
    batch.solve()
    for b in range(len(members)):
        k, x = batch.eigenpair(b)
        assert k == mix[b][0] and np.array_equal(x.get(), mix[b][1])
    # alone: a batch of one gives the same bits as the member in the mix
    for b in (0, 2, 6):
        alone = pk.Batch([Rs[b]], [Fs[b]], precision=prec)
        alone.solve()
        k, x = alone.eigenpair(0)
        assert k == mix[b][0], members[b][0]
        assert np.array_equal(x.get(), mix[b][1]), members[b][0]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_outer_acceleration_keeps_the_eigenpair(pk, members, precision):
    """Anderson mixing of the outer iterates (depth 1 to 3) against the plain
    update (depth 0): every member still meets the oracle tolerances, the mixed
    iteration needs fewer outer updates in all, and a depth outside 0..3 is
    refused."""
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    Rs, Fs = _mats(pk, members)
    outer = {}
    for depth in (0, 1, 3):
        print(f"[case] batch of {len(Rs)} {precision}, Anderson depth {depth}", flush=True)
        batch = pk.Batch(Rs, Fs, precision=prec, acceleration=depth)
        batch.solve()
        outer[depth] = 0
        for b, (name, G, ref) in enumerate(members):
            k, x = batch.eigenpair(b)
            st = batch.stats(b)
            dk = abs(k - ref.keff) * 1e5
            l2 = np.linalg.norm(_shape(x.get()) - _shape(ref.phi)) / np.linalg.norm(_shape(ref.phi))
            print(f"   {name}: dk={dk:.2e} pcm l2={l2:.2e} outer={st['outer']} inner={st['inner']}", flush=True)
            assert dk <= DK_PCM, (name, depth)
            assert l2 <= FLUX_L2, (name, depth)
            outer[depth] += st["outer"]
    assert outer[3] < outer[0] and outer[1] < outer[0], outer
    with pytest.raises(pk.PkError):
        pk.Batch(Rs[:1], Fs[:1], acceleration=4)


def test_batch_matches_single_eigen(pk, members):
    """The batch engine and pk_eig agree to the parity tolerances."""
    Rs, Fs = _mats(pk, members[:3])
    batch = pk.Batch(Rs, Fs)
    batch.solve()
    for b in range(3):
        kb, xb = batch.eigenpair(b)
        e = pk.Eigen(Rs[b], Fs[b])
        e.solve()
        ke, xe = e.eigenpair(Rs[b].size()[0])
        assert abs(kb - ke) * 1e5 <= DK_PCM
        sb, se = _shape(xb.get()), _shape(xe.get())
        assert np.linalg.norm(sb - se) / np.linalg.norm(se) <= FLUX_L2


def test_batch_initial_vector(pk, members):
    """Warm start from the converged vector: converges at the first check."""
    Rs, Fs = _mats(pk, members[:2])
    batch = pk.Batch(Rs, Fs)
    batch.solve()
    k0, x0 = batch.eigenpair(1)
    batch.set_initial(1, x0)
    batch.solve()
    k1, _ = batch.eigenpair(1)
    assert batch.stats(1)["outer"] <= 1
    assert abs(k1 - k0) * 1e5 <= DK_PCM


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pool_continuous_batching(pk, members, precision):
    """Members arrive in order into a pool that holds only a few at a time;
    each is admitted when a slot and memory free up, while the others iterate.
    Every result is bit-identical to the same member's in the fixed batch."""
    if precision not in _MIX:
        pytest.skip("needs the fixed-batch results of the same precision")
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    Rs, Fs = _mats(pk, members)
    pool = pk.Pool(max_rows=70000, max_entries=1 << 22, max_fentries=1 << 21, max_binv=1 << 21,
                   max_slots=3, precision=prec)
    queue = list(range(len(members)))
    slot_of, got = {}, {}
    waits = 0
    while queue or slot_of:
        while queue:
            s = pool.admit(Rs[queue[0]], Fs[queue[0]])
            if s is None:
                waits += 1
                break
            slot_of[s] = queue.pop(0)
        live, done = pool.step(1)
        for s in pool.collect():
            r = pool.result(s)
            # the batch interface reports lambda = 1/keff: compare through it
            got[slot_of.pop(s)] = (1.0 / (1.0 / r["keff"]), r["x"].get(), r["converged"])
            pool.release(s)
    assert waits > 0  # admission really waited for room
    for b, (k, x, conv) in got.items():
        assert conv, members[b][0]
        assert k == _MIX[precision][b][0], members[b][0]
        assert np.array_equal(x, _MIX[precision][b][1]), members[b][0]
    print(f"   pool: {len(got)} solves, {waits} admission waits, ticks={pool.capacity()['ticks']}",
          flush=True)


FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batched_classes.npz")


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="oracle fixture of the largest classes not generated")
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_largest_size_classes_against_oracle_fixture(pk, tmp_path, precision):
    """The batched family's four largest size classes (128x128 and 192x192 at 16
    groups, 256x256 at 8 and 16 groups: 262,144 to 1,048,576 rows) served
    together in one pool, against the FP64 oracle's keff and every 64th entry
    of its power-normalised flux (tests/golden/make_fixtures.py)."""
    from pampa_amd import workloads as W
    fx = np.load(FIXTURE)
    insts = sorted(int(k[4:-5]) for k in fx.files if k.endswith("_keff"))
    st = int(fx["stride"])
    paths = []
    for i in insts:
        d = str(tmp_path / f"inst-{i}")
        W.batched_instance(d, i)
        paths.append(os.path.join(d, "input.pmp"))
    prec = pk.PK_PREC_MIXED if precision == "mixed" else pk.PK_PREC_FP64
    srv = pk.Server(budget_bytes=64e9, max_slots=len(paths), precision=prec)
    for p in paths:
        srv.add_case(p)
    n = len(paths)
    res = srv.run(np.zeros(n), np.arange(n), np.zeros(n, np.uint64), rel=0.0, warm=False, keep_flux=list(range(n)))
    srv.close()
    for r, i in enumerate(insts):
        dk = abs(res["keff"][r] - float(fx[f"inst{i}_keff"])) * 1e5
        xs = np.asarray(res["flux"][r])[::st]
        xo = fx[f"inst{i}_phi"]
        l2 = np.linalg.norm(_shape(xs) - _shape(xo)) / np.linalg.norm(_shape(xo))
        print(f"[case] instance {i} {precision}: dk={dk:.2e} pcm flux L2 (every {st}th entry) {l2:.2e}", flush=True)
        assert dk <= DK_PCM
        assert l2 <= FLUX_L2


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_slice_heights_give_the_same_bits(pk, members, monkeypatch, precision):
    """The batch laid out with 32-, 64- and 128-row slices (PK_SLICE forces the
    autotuner's choice) gives bit-identical eigenpairs: padding entries add
    exact zeros and every row sums its entries in the same order."""
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    res = {}
    for h in (64, 32, 128):
        monkeypatch.setenv("PK_SLICE", str(h))
        print(f"[case] batch of {len(members)} {precision}, slice height {h}", flush=True)
        Rs, Fs = _mats(pk, members)
        batch = pk.Batch(Rs, Fs, precision=prec)
        batch.solve()
        res[h] = [(k, x.get()) for k, x in (batch.eigenpair(b) for b in range(len(members)))]
    for h in (32, 128):
        for b, (name, _, _) in enumerate(members):
            assert res[h][b][0] == res[64][b][0], (h, name)
            assert np.array_equal(res[h][b][1], res[64][b][1]), (h, name)


def test_autotuner_choices(pk, members, monkeypatch):
    """The tuner times every (slice height, workgroup) candidate once per size
    class; the height it keeps is no slower than 64 rows in 256-row
    workgroups (the pool's shape), and the workgroup no slower than 256 rows
    at that height.  A second batch of the same classes reuses the choices
    (no new classes) and gives the same bits as the first."""
    monkeypatch.delenv("PK_SLICE", raising=False)
    monkeypatch.delenv("PK_TUNE", raising=False)
    pk.tune_reset()
    Rs, Fs = _mats(pk, members)
    batch = pk.Batch(Rs, Fs, precision=pk.PK_PREC_MIXED)
    batch.solve()
    first = [(k, x.get()) for k, x in (batch.eigenpair(b) for b in range(len(members)))]
    ch = pk.tune_choices()
    classes = {(r.size()[0], r.size()[1], G) for r, (_, G, _) in zip(Rs, members)}
    assert len(ch) == len(classes)
    for c in ch:
        print(f"   rows {c['rows']} G {c['block']}: slice height {c['slice_height']}, SpMV workgroup "
              f"{c['spmv_block']}", flush=True)
        assert c["slice_height"] in (32, 64, 128) and c["spmv_block"] in (128, 256, 512, 1024)
        h, blk, us = c["slice_height"], c["spmv_block"], c["us"]
        assert all(us[f"64x{b}"] > 0 for b in (128, 256, 512, 1024))
        assert 0 < us[f"{h}x256"] <= us["64x256"]
        assert us[f"{h}x{blk}"] <= us[f"{h}x256"]
    Rs2, Fs2 = _mats(pk, members)
    batch2 = pk.Batch(Rs2, Fs2, precision=pk.PK_PREC_MIXED)
    batch2.solve()
    assert len(pk.tune_choices()) == len(classes)
    for b in range(len(members)):
        k, x = batch2.eigenpair(b)
        assert k == first[b][0] and np.array_equal(x.get(), first[b][1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tick_length_gives_the_same_bits(pk, members, monkeypatch, precision):
    """A GMRES cycle may span ticks: each problem keeps its step on the device
    and the tick's end finishes only the cycles that stopped.  Ticks of 1, 4
    and 13 step launches (PK_TICK_STEPS; cycles then span several ticks, in a
    batch with restart 60 and 16-group members capped at 30) give the same
    eigenpairs, bit for bit, as the predicted tick lengths."""
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    res = {}
    for n in (0, 13, 4, 1):
        if n:
            monkeypatch.setenv("PK_TICK_STEPS", str(n))
        else:
            monkeypatch.delenv("PK_TICK_STEPS", raising=False)
        print(f"[case] batch of {len(members)} {precision}, tick steps {n or 'predicted'}", flush=True)
        Rs, Fs = _mats(pk, members)
        batch = pk.Batch(Rs, Fs, precision=prec)
        batch.solve()
        res[n] = [(k, x.get(), batch.stats(b)) for b, (k, x) in
                  enumerate(batch.eigenpair(b) for b in range(len(members)))]
    for n in (13, 4, 1):
        for b, (name, _, _) in enumerate(members):
            assert res[n][b][0] == res[0][b][0], (n, name)
            assert np.array_equal(res[n][b][1], res[0][b][1]), (n, name)
            assert res[n][b][2]["outer"] == res[0][b][2]["outer"], (n, name)
            assert res[n][b][2]["inner"] == res[0][b][2]["inner"], (n, name)
