"""Serving (pampa_serve_*): requests on resident base cases; diffusion
requests continuously batched into one device eigen pool, S_N requests solved
by per-case workers beside it.

  * the perturbed cross sections a request solves with are bit-identical to
    the published perturbation rule (pampa_amd.workloads.perturb_tables)
    applied to the oracle's tables of the same input;
  * every sampled request matches the FP64 oracle's solve of the same perturbed
    problem within |dk| <= 1 pcm and flux relative L2 <= 1e-5, cold and warm;
  * a request's result is bit-identical alone and in the mix.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DK_PCM = 1.0
FLUX_L2 = 1e-5
REL = 0.02


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    from pampa_amd import workloads as W
    base = tmp_path_factory.mktemp("serve")
    out = []
    for name, idx in (("slab-1d", None), ("inst-9", 9), ("inst-5", 5), ("inst-0", 0)):
        d = str(base / name)
        if idx is None:
            W.slab_1d(d)
        else:
            W.batched_instance(d, idx)
        out.append((name, os.path.join(d, "input.pmp")))
    return out


def _oracle(oracle_mod, path, seed, rel):
    from pampa_amd import workloads as W
    case = oracle_mod.pmp.read_input(path)
    mesh = oracle_mod.build_mesh(case)
    s = case.solver
    xs = oracle_mod.xs_tables(case.materials, s.G, diffusion=True)
    xs.sigt, xs.sigs, xs.D, xs.nusf = W.perturb_tables(xs.sigt, xs.sigs, xs.D, xs.nusf, seed, rel)
    bct, bcv = oracle_mod.bc_table(mesh, s.bcs)
    R, F = oracle_mod.assemble_diffusion(mesh, xs, bct, bcv)
    k, x = oracle_mod.eigen_sinvert(R, F)
    return k, x, xs


def _shape(v):
    v = np.asarray(v, dtype=float)
    return v / v.sum()


def _stream(ncases, n, seed0=1000):
    rng = np.random.default_rng(7)
    cid = rng.integers(0, ncases, n)
    seeds = np.arange(seed0, seed0 + n, dtype=np.uint64)
    return np.zeros(n), cid, seeds


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_serve_parity_and_alone_vs_mix(pk, oracle_mod, cases, precision):
    prec = pk.PK_PREC_MIXED if precision == "mixed" else pk.PK_PREC_FP64
    srv = pk.Server(budget_bytes=2e9, max_slots=4, precision=prec)
    ids = [srv.add_case(p)[0] for _, p in cases]
    n = 16
    arrival, cid, seeds = _stream(len(ids), n)
    print(f"[case] serve {n} requests over {len(ids)} cases, {precision}", flush=True)
    # cold stream (burst: every request waits for room in a 4-slot pool)
    cold = srv.run(arrival, cid, seeds, rel=REL, warm=False, keep_flux=list(range(n)))
    assert cold["converged"].all()
    # tables of the requests: bit-identical to the published rule
    for r in (0, 5):
        k_o, x_o, xs = _oracle(oracle_mod, cases[cid[r]][1], int(seeds[r]), REL)
        nm, G = xs.sigt.shape
        st, ss, D, nf = srv.case_tables(int(cid[r]), nm, G, int(seeds[r]), REL)
        assert np.array_equal(st, xs.sigt) and np.array_equal(ss, xs.sigs)
        assert np.array_equal(D, xs.D) and np.array_equal(nf, xs.nusf)
    # parity of a sample against the oracle
    for r in range(0, n, 3):
        k_o, x_o, _ = _oracle(oracle_mod, cases[cid[r]][1], int(seeds[r]), REL)
        dk = abs(cold["keff"][r] - k_o) * 1e5
        l2 = np.linalg.norm(_shape(cold["flux"][r]) - _shape(x_o)) / np.linalg.norm(_shape(x_o))
        print(f"   req {r} case {cases[cid[r]][0]}: dk={dk:.2e} pcm l2={l2:.2e} "
              f"outer={cold['outer'][r]} inner={cold['inner'][r]}", flush=True)
        assert dk <= DK_PCM and l2 <= FLUX_L2
    # alone: the same requests one at a time give the same bits
    for r in (1, 4, 9):
        one = srv.run([0.0], [cid[r]], [seeds[r]], rel=REL, warm=False, keep_flux=[0])
        assert one["keff"][0] == cold["keff"][r]
        assert np.array_equal(one["flux"][0], cold["flux"][r])
    # warm stream: requests start from their case's latest converged flux
    srv.run(np.zeros(len(ids)), np.arange(len(ids)), np.zeros(len(ids), np.uint64), rel=REL, warm=False)
    warm = srv.run(arrival, cid, seeds + np.uint64(500), rel=REL, warm=True, keep_flux=[2, 7])
    assert warm["converged"].all()
    assert warm["outer"].sum() < cold["outer"].sum()
    for r in (2, 7):
        k_o, x_o, _ = _oracle(oracle_mod, cases[cid[r]][1], int(seeds[r]) + 500, REL)
        assert abs(warm["keff"][r] - k_o) * 1e5 <= DK_PCM
        l2 = np.linalg.norm(_shape(warm["flux"][r]) - _shape(x_o)) / np.linalg.norm(_shape(x_o))
        assert l2 <= FLUX_L2
    st = srv.stats()
    print(f"   warm outer {warm['outer'].sum()} vs cold {cold['outer'].sum()}; ticks={st['ticks']} "
          f"max_live={st['max_live']}", flush=True)
    srv.close()


def test_serve_open_loop_arrivals(pk, cases):
    """Requests that arrive while others run are admitted as slots free up."""
    srv = pk.Server(budget_bytes=2e9, max_slots=8, precision=pk.PK_PREC_MIXED)
    ids = [srv.add_case(p)[0] for _, p in cases[:2]]
    n = 40
    arrival = np.arange(n) * 0.002  # 500 req/s
    cid = np.arange(n) % len(ids)
    res = srv.run(arrival, cid, np.arange(1, n + 1, dtype=np.uint64), rel=REL, warm=True)
    assert res["converged"].all()
    assert (res["latency"] > 0).all()
    assert srv.stats()["max_live"] > 1
    srv.close()


def _oracle_sn(oracle_mod, path, seed, rel):
    from pampa_amd import workloads as W
    case = oracle_mod.pmp.read_input(path)
    mesh = oracle_mod.build_mesh(case)
    s = case.solver
    xs = oracle_mod.xs_tables(case.materials, s.G, diffusion=False)
    xs.sigt, xs.sigs, _, xs.nusf = W.perturb_tables(xs.sigt, xs.sigs, None, xs.nusf, seed, rel)
    k, _, phiw, _ = oracle_mod.sn_sweep_eigen(mesh, xs, s, inner=1, tol_k=1e-12, tol_phi=1e-11)
    return k, phiw


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_serve_sn_requests(pk, oracle_mod, cases, tmp_path, precision):
    """S_N requests (hexagonal lattices, S4 and S6) served beside diffusion
    requests: each matches the FP64 oracle's sweep solve of the same perturbed
    problem (|dk| <= 1 pcm, moment relative L2 <= 1e-5), cold and warm, with
    one and with two workers per case; a cold request's result is
    bit-identical alone and in the mix (its solve runs on its worker's own
    stream, whatever else runs)."""
    from pampa_amd import workloads as W
    sn_paths = []
    for j, (layers, order) in enumerate(((2, 4), (3, 6))):
        d = str(tmp_path / f"sn-{j}")
        W.hex_sn_3d(d, seed=100 + j, rings=2, layers=layers, order=order)
        sn_paths.append(os.path.join(d, "input.pmp"))
    prec = pk.PK_PREC_MIXED if precision == "mixed" else pk.PK_PREC_FP64
    srv = pk.Server(budget_bytes=2e9, max_slots=4, precision=prec)
    srv.sn_workers(2 if precision == "mixed" else 1)
    dif = [srv.add_case(p)[0] for _, p in cases[:2]]
    sn = [srv.add_case(p)[0] for p in sn_paths]
    n = 12
    cid = np.array([sn[0], dif[0], sn[1], dif[1], sn[0], sn[1], dif[0], sn[0], dif[1], sn[1], sn[0], dif[0]])
    seeds = np.arange(2000, 2000 + n, dtype=np.uint64)
    arrival = np.arange(n) * 0.001
    print(f"[case] serve {n} requests, S_N and diffusion, {precision}", flush=True)
    cold = srv.run(arrival, cid, seeds, rel=REL, warm=False, keep_flux=list(range(n)))
    assert cold["converged"].all()
    for r in range(n):
        if cid[r] not in sn:
            continue
        k_o, phiw_o = _oracle_sn(oracle_mod, sn_paths[sn.index(cid[r])], int(seeds[r]), REL)
        dk = abs(cold["keff"][r] - k_o) * 1e5
        l2 = np.linalg.norm(_shape(cold["flux"][r]) - _shape(phiw_o)) / np.linalg.norm(_shape(phiw_o))
        print(f"   req {r} S_N case {cid[r]}: dk={dk:.2e} pcm l2={l2:.2e} outer={cold['outer'][r]} "
              f"sweeps={cold['inner'][r]}", flush=True)
        assert dk <= DK_PCM and l2 <= FLUX_L2
    # alone: an S_N request replayed by itself (cold) gives the same bits
    for r in (2, 4):
        one = srv.run([0.0], [cid[r]], [seeds[r]], rel=REL, warm=False, keep_flux=[0])
        assert one["keff"][0] == cold["keff"][r]
        assert np.array_equal(one["flux"][0], cold["flux"][r])
    # warm: S_N requests start from their case's previous solution
    warm = srv.run(arrival, cid, seeds + np.uint64(50), rel=REL, warm=True, keep_flux=[0, 5])
    assert warm["converged"].all()
    snr = np.isin(cid, sn)
    assert warm["outer"][snr].sum() < cold["outer"][snr].sum()
    for r in (0, 5):
        k_o, phiw_o = _oracle_sn(oracle_mod, sn_paths[sn.index(cid[r])], int(seeds[r]) + 50, REL)
        assert abs(warm["keff"][r] - k_o) * 1e5 <= DK_PCM
        l2 = np.linalg.norm(_shape(warm["flux"][r]) - _shape(phiw_o)) / np.linalg.norm(_shape(phiw_o))
        assert l2 <= FLUX_L2
    srv.close()
