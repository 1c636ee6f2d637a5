"""Mixed precision with FP64 refinement across condition numbers, at full size (C4).

The random Zipf factor (n = 1e6) and the 4,096-supernode factor with the diagonal scaled so that
kappa_1 spans 1e2 ... 1e8 (gen.KAPPA_SCALES, calibrated with tests/golden/calibrate_kappa.py).
Inner solves: FP32 (Zipf), BF16 rectangular products with FP32 accumulation (supernodal).  Bar:
the refined x within 1e-12 of the serial oracle, reached by converging or, where the low
precision stagnates, by the FP64 fallback (the outcome is checked either way); the kappa
estimate on the device within a factor 3 of the target.  The supernodal FP64 solve, which
multiplies by explicit inverses of the diagonal blocks, is checked at kappa 1e8 as well."""
import numpy as np
import pytest

from oracle import oracle as O
from sptrsv_amd import capi, gen

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
_factors = {}


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def oracle_rows(U, b):
    if b.ndim == 1:
        return O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0]
    x = O.base_gauss_upper_columns(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))
    return np.ascontiguousarray(x.T)


def plan_for(fam, kappa):
    """one resident plan at a time (the supernodal factor holds 82M entries)"""
    key = (fam, kappa)
    if key not in _factors:
        for k in list(_factors):
            _factors.pop(k)[1].close()
        U = gen.kappa_scaled(fam, kappa)
        banner(f"plan {fam} kappa {kappa:.0e}")
        P = capi.Plan.from_upper(U)
        if fam.startswith("supernodal"):
            P.build_supernodes(U.snodes)
        _factors[key] = (U, P)
    return _factors[key]


@pytest.fixture(scope="module", autouse=True)
def _release():
    yield
    for k in list(_factors):
        _factors.pop(k)[1].close()


CASES = [("zipf_1e6", 1e2, 1), ("zipf_1e6", 1e8, 1), ("zipf_1e6", 1e8, 32),
         ("supernodal_4096", 1e2, 32), ("supernodal_4096", 1e8, 1), ("supernodal_4096", 1e8, 32),
         ("supernodal_4096", 1e8, 128)]


@pytest.mark.parametrize("fam,kappa,m", CASES)
def test_refined_full_size(fam, kappa, m):
    U, P = plan_for(fam, kappa)
    sn = fam.startswith("supernodal")
    banner(f"refine {fam} kappa {kappa:.0e} m={m}")
    xt, b = gen.rhs_for(U, m, seed=5, layout="row")
    d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
    prm = capi.params(variant=capi.VAR_SUPERNODAL if sn else capi.VAR_AUTO)
    info = P.solve_refined_ex(d_b, d_x, m, prm, precision=capi.BF16 if sn else capi.FP32, fallback=True)
    x = d_x.download()
    d_b.free()
    d_x.free()
    assert info["outcome"] == "converged" or info["fell_back"], info
    err = O.rel_error(oracle_rows(U, b), x)
    banner(f"refine {fam} kappa {kappa:.0e} m={m}: {info}, rel err {err:.2e}")
    assert err <= FP64_TOL


@pytest.mark.parametrize("fam,kappa", [("zipf_1e6", 1e8), ("supernodal_4096", 1e8)])
def test_kappa_estimate(fam, kappa):
    U, P = plan_for(fam, kappa)
    k1, _ = P.condest()
    banner(f"condest {fam} target {kappa:.0e}: {k1:.3e}")
    assert kappa / 3 <= k1 <= kappa * 3


def test_supernodal_explicit_inverse_kappa_1e8():
    """FP64 supernodal solve (explicit inverses of the diagonal blocks) at kappa 1e8"""
    U, P = plan_for("supernodal_4096", 1e8)
    xt, b = gen.rhs_for(U, 8, seed=6, layout="row")
    x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
    assert O.rel_error(oracle_rows(U, b), x) <= FP64_TOL
