"""The learned variant selector of single systems (AUTO): the shipped network
(sptrsv_amd/selector_single.json, trained by `python -m sptrsv_amd.tune_single`) judged on the
held-out factors' timing table it was evaluated on (tests/golden/selector_single_heldout.json:
features, wall ms of every valid variant, the hand rule's pick), and -- on the GPU -- the
library's own forward pass and AUTO against numpy's."""
import json
import os

import numpy as np
import pytest

from sptrsv_amd import capi, gen, tune_single

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NET = os.path.join(ROOT, "sptrsv_amd", "selector_single.json")
HELDOUT = os.path.join(ROOT, "tests", "golden", "selector_single_heldout.json")


def load():
    with open(NET) as f:
        doc = json.load(f)
    with open(HELDOUT) as f:
        h = json.load(f)
    T = np.asarray(h["times_ms"], dtype=np.float64)
    T[T < 0] = np.inf
    return doc, np.asarray(h["features"], dtype=np.float32), T, np.asarray(h["hand_rule"]), h


def test_network_file_layout():
    doc, F, T, H, h = load()
    i, h1, h2, o = doc["sizes"]
    assert (i, o) == (24, len(doc["classes"])) and doc["classes"] == tune_single.CLASSES == h["classes"]
    assert len(doc["params"]) == i * h1 + h1 + h1 * h2 + h2 + h2 * o + o
    assert len(doc["mu"]) == len(doc["sd"]) == 24 and min(doc["sd"]) > 0
    assert F.shape == (len(T), 24) and T.shape[1] == o


def test_selector_beats_fallbacks_on_heldout_factors():
    """held-out factors (another seed than training): the network's picks cost at most a few
    percent over exhaustive search, less than the plain sync-free kernel (the reference's
    algorithm) by far, and no more than the hand rule by more than a percent"""
    doc, F, T, H, h = load()
    pick = tune_single.predict(doc, F, T)
    assert np.all(np.isfinite(T[np.arange(len(T)), pick])), "a pick outside the valid variants"
    r_mlp = tune_single.regret(T, pick)
    r_hand = tune_single.regret(T, np.maximum(H, 0))
    r_push = tune_single.regret(T, np.full(len(T), tune_single.CLASSES.index(capi.VAR_CSC_PUSH)))
    assert r_mlp == pytest.approx(doc["meta"]["test"]["regret_mlp"], rel=1e-6, abs=1e-9)
    assert r_mlp <= 0.10
    assert r_mlp <= r_hand + 0.01
    assert r_push > 10 * max(r_mlp, 0.01)


def test_selector_set_validates_its_inputs():
    """host-only entry points (no device call): a network with a bad variant id, a zero or
    non-finite feature scale, or wrong input width is rejected; a valid one installs and clears"""
    doc, *_ = load()
    sizes, params, mu, sd, classes = doc["sizes"], doc["params"], doc["mu"], doc["sd"], doc["classes"]
    bad_sd = list(sd)
    bad_sd[3] = 0.0
    nan_mu = list(mu)
    nan_mu[0] = float("nan")
    for args in ((sizes, params, mu, sd, [0] + classes[1:]), (sizes, params, mu, bad_sd, classes),
                 (sizes, params, nan_mu, sd, classes), ([23] + sizes[1:], params, mu, sd, classes)):
        with pytest.raises(capi.SptrsvError):
            capi.selector_set(*args)
    capi.selector_set(sizes, params, mu, sd, classes)
    capi.selector_clear()


@pytest.mark.gpu
def test_library_forward_and_auto_match_numpy():
    """the library's forward pass on a plan's features picks the variant numpy's forward pass
    scores highest among the valid ones; AUTO runs it and solves to the oracle; clearing the
    network brings the hand rule back"""
    from oracle import oracle as O
    doc, *_ = load()
    cases = [("grid3d", tune_single.grid_factor([20, 18, 16], 1)), ("zipf", gen.random_zipf(20000, seed=3)),
             ("chol2d", gen.laplace2d_cholesky(40)), ("banded", gen.banded(30000, 9, seed=2)),
             ("supernodal", gen.random_supernodal(40, seed=4, wmin=8, wmax=64))]
    try:
        for fam, U in cases:
            with capi.Plan.from_upper(U) as P:
                if fam == "supernodal":
                    P.build_supernodes(U.snodes)
                # the variants valid here, decided as the autotuner's table decided them
                capi.selector_clear()
                has_sn = P.auto_variant(1) == capi.VAR_SUPERNODAL
                grid = P.auto_variant(1) == capi.VAR_BLOCK or P.auto_variant(40) == capi.VAR_PANEL
                capi.selector_load(NET)
                for m in (1, 4, 32):
                    ok = [c for c in tune_single.CLASSES if tune_single.valid(P, c, m, grid, has_sn)]
                    v = P.auto_variant(m)
                    f = P.features(m)
                    s = tune_single.forward(doc, f[None, :])[0]
                    k = tune_single.CLASSES.index(v)
                    # the library's choice is the best valid score (float32 vs float64: near ties)
                    assert v in ok, (fam, m, v)
                    assert s[k] >= np.max(s[[tune_single.CLASSES.index(c) for c in ok]]) - 1e-3 * (1 + abs(s[k]))
                    xt, b = gen.rhs_for(U, m, seed=5, layout="row")
                    x = P.solve_host(b, capi.params())
                    bb = b if m == 1 else np.ascontiguousarray(b.T)
                    ref = O.base_gauss_upper(U.rowptr, U.col, U.val, bb)[0]
                    assert O.rel_error(ref if m == 1 else ref.T, x) <= 1e-12, (fam, m, v)
    finally:
        capi.selector_clear()
    U = gen.random_zipf(5000, seed=3)
    with capi.Plan.from_upper(U) as P:
        assert P.auto_variant(2) == capi.VAR_LEVEL_PULL  # the hand rule again

