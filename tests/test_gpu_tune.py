"""The autotuner: every launch shape it can choose gives the same results, the
tuner picks one of them, and its database round-trips through a file."""
import os

import numpy as np
import pytest

import oracle
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
def test_every_plan_shape_matches_oracle(gpu, prec, tol):
    k = 1
    B = W.torus_blocks(6, 6, seed=4, k0=1, ns_range=(64, 900), nt_range=(1, 260))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    ref = np.zeros(nt)
    for b in range(B["trg_cnt"].size):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        ref[t0:t0 + tc] = oracle.eval(k, B["r_trg"][3 * t0:3 * (t0 + tc)], B["r_src"][3 * s0:3 * (s0 + sc)],
                                      B["n_src"][3 * s0:3 * (s0 + sc)], B["v_src"][s0:s0 + sc])
    P = gpu.Plan(k, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"])
    for spl in (1, 2, 4):
        for tpl in (1, 2):
            for cap in (0, 4, 3, 2, 1):
                P.set_shape(spl, tpl, cap)
                assert P.shape() == (spl, tpl, cap)
                out = P.apply(B["v_src"], np.zeros(nt))
                assert W.rel_errors(out, ref)[0] < tol, (spl, tpl, cap)


def test_plan_tune_reports_the_configured_shape(gpu, tmp_path):
    """A kernel with nine outputs per target runs one target per lane whatever
    is asked: the tuner reports (and stores) the shape it actually ran."""
    gpu.tune_db_path(str(tmp_path / "tune9.db"))
    try:
        B = W.torus_blocks(8, 8, seed=1, k0=1)
        nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
        P = gpu.Plan(6, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=1)
        P.set_geometry(B["r_trg"], B["r_src"], None)
        v = gpu.DeviceArray.from_host(B["v_src"])
        scratch = gpu.DeviceArray(nt * 9)
        spl, tpl, cap, ms = P.tune(v, scratch)
        assert tpl == 1 and ms > 0
        assert P.shape() == (spl, tpl, cap)
    finally:
        gpu.tune_db_path(None)


def test_plan_tune_picks_a_candidate(gpu, tmp_path):
    gpu.tune_db_path(str(tmp_path / "tune.db"))
    try:
        B = W.torus_blocks(16, 16, seed=0, k0=3)
        nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
        P = gpu.Plan(3, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=1)
        P.set_geometry(B["r_trg"], B["r_src"], B["n_src"])
        v = gpu.DeviceArray.from_host(B["v_src"])
        scratch = gpu.DeviceArray(nt * 3)
        spl, tpl, cap, ms = P.tune(v, scratch)
        assert spl in (1, 2, 4) and tpl in (1, 2) and cap in (0, 4, 3, 2, 1) and ms > 0
        assert P.shape() == (spl, tpl, cap)
        text = open(tmp_path / "tune.db").read()
        assert "plan k3 p1" in text
        # a new plan with the same size class starts from the tuned shape
        Q = gpu.Plan(3, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=1)
        assert Q.shape() == (spl, tpl, cap)
        # and still computes the right thing
        Q.set_geometry(B["r_trg"], B["r_src"], B["n_src"])
        a = Q.apply(B["v_src"], np.zeros(nt * 3))
        P.set_shape(1, 1, 0)
        b = P.apply(B["v_src"], np.zeros(nt * 3))
        assert W.rel_errors(a, b)[0] < 1e-6  # mixed precision: shapes round their FP32 partials differently
    finally:
        gpu.tune_db_path(None)


def _log2(n):
    return int(n).bit_length() - 1


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
def test_dense_shapes_from_database_match_oracle(gpu, tmp_path, prec, tol):
    rng = np.random.default_rng(8)
    nt, ns = 3000, 2500
    xt, xs, n = rng.random(nt * 3), rng.random(ns * 3), rng.standard_normal(ns * 3)
    v = rng.uniform(-1, 1, ns * 3)
    ref = oracle.eval(2, xt, xs, None, v)
    db = tmp_path / "dense.db"
    newton = 2 if prec == 0 else 0
    try:
        for tpl in (1, 2, 4):
            for cap in (0, 3, 2):
                db.write_text("dense k2 p%d n%d t%d s%d | %d %d 1.0\n" % (prec, newton, _log2(nt), _log2(ns), tpl, cap))
                gpu.tune_db_path(str(db))
                out = gpu.eval(2, xt, xs, None, v, precision=prec)
                assert W.rel_errors(out, ref)[0] < tol, (tpl, cap)
    finally:
        gpu.tune_db_path(None)


@pytest.mark.parametrize("k", range(16))
def test_mixed_every_targets_per_lane_matches_oracle(gpu, tmp_path, k):
    """Mixed precision at each targets-per-lane count 1, 2, 4 (a kernel with
    more outputs per target caps it lower) for every micro-kernel, on
    uneven sizes with coincident points.  FP32 pairwise terms are accurate
    relative to their own size, so the error is bounded relative to the sum
    of the terms' magnitudes, |K|^T |f| from the oracle's kernel matrix."""
    rng = np.random.default_rng(40 + k)
    _, K0, _, _, _ = oracle.dims(k)
    nt, ns = 801, 703
    xs, xt = rng.random(ns * 3), rng.random(nt * 3)
    xt[:90] = xs[:90]
    n = rng.standard_normal(ns * 3)
    n = (n.reshape(-1, 3) / np.linalg.norm(n.reshape(-1, 3), axis=1)[:, None]).ravel()
    v = rng.uniform(-1, 1, ns * K0)
    mu = 1.0 if k >= 12 else None
    ref = oracle.eval(k, xt, xs, n, v, ctx=mu)
    magnitude = np.max(np.abs(oracle.kernel_matrix(k, xt, xs, n, ctx=mu)).T @ np.abs(v))
    gpu.tune_db_path(str(tmp_path / "tpl.db"))
    try:
        for tpl in (1, 2, 4):
            gpu.set_dense_shape(k, nt, ns, tpl, 0, precision=gpu.MIXED)
            out = gpu.eval(k, xt, xs, n, v, precision=gpu.MIXED, ctx=mu)
            assert np.isfinite(out).all()
            e = np.max(np.abs(out - ref)) / magnitude
            assert e < 1e-6, (tpl, e)
    finally:
        gpu.tune_db_path(None)


def test_tune_eval_records_choice(gpu, tmp_path):
    gpu.tune_db_path(str(tmp_path / "t.db"))
    try:
        xt, xs, f = W.cube(40000, 30000, seed=2)
        d = [gpu.DeviceArray.from_host(a) for a in (xt, xs, f)]
        scratch = gpu.DeviceArray(40000)
        tpl, cap, ms = gpu.tune_eval(0, d[0], d[1], None, d[2], scratch, precision=1)
        assert tpl in (1, 2, 4) and cap in (0, 3, 2) and ms > 0
        assert gpu.dense_shape(0, 30000, 40000, precision=1) == (tpl, cap)  # cube(n_src, n_trg)
        assert "dense k0 p1 n0" in open(tmp_path / "t.db").read()
        ref = oracle.eval(0, xt[:3000], xs, None, f)
        out = gpu.eval(0, xt, xs, None, f, precision=1)
        assert W.rel_errors(out[:1000], ref)[0] < 1e-6
    finally:
        gpu.tune_db_path(None)
