"""HIP QG kernels against the FP64 oracle (through the C ABI).

Bars: FP64 plans agree with the oracle to 1e-10 relative (the DST is a GEMM
here and an FFT in the oracle, so only rounding differs); FP32 plans to the
tolerances written in each test (FP32 rounding over a few steps).
"""
import numpy as np
import pytest
import torch

from oracle.qg import QgmOracle
from vardyn_amd.qg import QgPlan, build_mask

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = ~(np.isnan(a) | np.isnan(b))
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN pattern differs"
    return np.linalg.norm(a[m] - b[m]) / max(np.linalg.norm(b[m]), 1e-300)


def make_pair(ny, nx, scheme, precision, land=False, beta=False, kdiff=False, mdt=False,
              bathy=False, seed=0, dst_dense=False, dst_split=0):
    rng = np.random.default_rng(seed)
    dx, dy, dt, c = 10e3, 12e3, 1200.0, 2.7
    ssh = None
    if land:
        ssh = np.zeros((ny, nx))
        ssh[ny // 2, nx // 3] = np.nan
        ssh[ny // 2 + 1, nx // 3] = np.nan
        ssh[3, nx - 6] = np.nan
    f = 1e-4
    if beta:
        f = 1e-4 + 2e-11 * dy * (np.arange(ny)[:, None] - ny / 2) * np.ones((1, nx))
    K = 40.0 if kdiff else None
    m = 0.05 * rng.standard_normal((ny, nx)) if mdt else None
    b = 0.1 * rng.standard_normal((ny, nx)) if bathy else None
    orc = QgmOracle(dx=dx, dy=dy, dt=dt, c=c, f=f, time_scheme=scheme, shape=(ny, nx), SSH=ssh,
                    Kdiffus=K, mdt=m, bathymetry_PV_term=b)
    fb = None if b is None else float(orc.f) * b
    plan = QgPlan(ny, nx, dx, dy, dt, c, float(orc.f), time_scheme=scheme, precision=precision,
                  kdiffus=K, mask=build_mask(ny, nx, ssh), beta=orc.beta, fbathy=fb, mdt=m,
                  dst_dense=dst_dense, dst_split=dst_split)
    h0 = 0.1 * rng.standard_normal((ny, nx))
    if land:
        h0[np.isnan(ssh)] = np.nan
    hb = 0.01 * rng.standard_normal((ny, nx))
    return orc, plan, h0, hb, rng


def test_mask_matches_oracle():
    from oracle.qg import build_mask as omask
    ssh = np.zeros((20, 17))
    ssh[5, 6] = np.nan
    ssh[12, 2] = np.nan
    assert np.array_equal(build_mask(20, 17, ssh), omask(20, 17, ssh))


@pytest.mark.parametrize("dst_dense", [False, True])
@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 2e-6)])
@pytest.mark.parametrize("shape", [(40, 48), (66, 34), (128, 128), (41, 47), (40, 45)])
def test_pv2h_h2pv(precision, tol, shape, dst_dense):
    ny, nx = shape
    orc, plan, h0, hb, rng = make_pair(ny, nx, "rk2", precision, dst_dense=dst_dense)
    q = rng.standard_normal((ny, nx)) * 1e-5
    qb = orc.h2pv(hb, hb)
    h_ref = orc.pv2h(q, hb, qb)
    h = plan.pv2h(plan.field(q), plan.field(hb), plan.field(qb))
    torch.cuda.synchronize()
    assert rel(h.cpu().numpy(), h_ref) < tol
    q_ref = orc.h2pv(h0, hb)
    qq = plan.h2pv(plan.field(h0), plan.field(hb))
    assert rel(qq.cpu().numpy(), q_ref) < tol


@pytest.mark.parametrize("dst_dense", [False, True])
@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
@pytest.mark.parametrize("precision,tol", [("float64", 1e-10), ("float32", 2e-5)])
def test_step(scheme, precision, tol, dst_dense):
    orc, plan, h0, hb, _ = make_pair(40, 48, scheme, precision, dst_dense=dst_dense)
    ref = orc.step(h0, hb, 6)
    h1 = plan.step(plan.field(h0), plan.field(hb), 6)
    assert rel(h1.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("opts", [dict(land=True), dict(beta=True), dict(kdiff=True),
                                  dict(mdt=True), dict(bathy=True),
                                  dict(land=True, beta=True, kdiff=True, mdt=True, bathy=True)])
@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_step_options_f64(scheme, opts):
    orc, plan, h0, hb, _ = make_pair(36, 44, scheme, "float64", **opts)
    ref = orc.step(h0, hb, 4)
    h1 = plan.step(plan.field(h0), plan.field(hb), 4)
    assert rel(h1.cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("opts", [dict(), dict(land=True, beta=True, kdiff=True, mdt=True,
                                               bathy=True)])
@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_tangent_adjoint_f64(scheme, opts):
    orc, plan, h0, hb, rng = make_pair(36, 44, scheme, "float64", **opts)
    dh = 0.1 * rng.standard_normal(h0.shape)
    ay = rng.standard_normal(h0.shape)
    n = 5
    tl_ref, _ = orc.step_tgl(dh, h0, hb, n)
    ad_ref = orc.step_adj(ay, h0, hb, n)
    tl = plan.step_tgl(plan.field(dh), plan.field(h0), plan.field(hb), n).cpu().numpy()
    ad = plan.step_adj(plan.field(ay), plan.field(h0), plan.field(hb), n).cpu().numpy()
    assert rel(tl, tl_ref) < 1e-10
    assert rel(ad, ad_ref) < 1e-10
    # dot-product identity on the GPU results themselves
    h1 = orc.step(h0, hb, n)
    m = ~np.isnan(h1)
    p1 = np.sum(tl[m] * ay[m])
    p2 = np.nansum(dh * ad)
    assert abs(p1 / p2 - 1) < 1e-11


@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_tangent_adjoint_f32(scheme):
    orc, plan, h0, hb, rng = make_pair(64, 64, scheme, "float32")
    dh = 0.1 * rng.standard_normal(h0.shape)
    ay = rng.standard_normal(h0.shape)
    n = 4
    tl_ref, _ = orc.step_tgl(dh, h0, hb, n)
    ad_ref = orc.step_adj(ay, h0, hb, n)
    tl = plan.step_tgl(plan.field(dh), plan.field(h0), plan.field(hb), n).cpu().numpy()
    ad = plan.step_adj(plan.field(ay), plan.field(h0), plan.field(hb), n).cpu().numpy()
    assert rel(tl, tl_ref) < 1e-4
    assert rel(ad, ad_ref) < 1e-4


def test_adjoint_tape_reuse_bitwise():
    """Adjoint from a tape written by step() equals the recomputing adjoint bit for bit."""
    orc, plan, h0, hb, rng = make_pair(48, 40, "rk2", "float64")
    n = 7
    H0, HB = plan.field(h0), plan.field(hb)
    ay = plan.field(rng.standard_normal(h0.shape))
    tape = plan.new_tape(n)
    plan.step(H0, HB, n, tape=tape)
    a1 = plan.step_adj(ay, H0, HB, n, tape=tape, tape_valid=True).clone()
    a2 = plan.step_adj(ay, H0, HB, n)
    assert torch.equal(a1, a2)


def test_nstep_zero():
    orc, plan, h0, hb, _ = make_pair(32, 32, "rk2", "float64", land=True)
    ref = orc.step(h0, hb, 0)
    h1 = plan.step(plan.field(h0), plan.field(hb), 0).cpu().numpy()
    assert rel(h1, ref) < 1e-14


@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_null_boundary_equals_zero_field(scheme):
    """hb = NULL (zero boundary) skips h_b/q_b and gives the same fields."""
    orc, plan, h0, hb, rng = make_pair(40, 48, scheme, "float64")
    H0, Z = plan.field(h0), plan.field(np.zeros_like(h0))
    ay = plan.field(rng.standard_normal(h0.shape))
    dh = plan.field(rng.standard_normal(h0.shape))
    assert torch.equal(plan.step(H0, None, 5), plan.step(H0, Z, 5))
    assert torch.equal(plan.step_tgl(dh, H0, None, 5), plan.step_tgl(dh, H0, Z, 5))
    assert torch.equal(plan.step_adj(ay, H0, None, 5), plan.step_adj(ay, H0, Z, 5))
    ref = orc.step(h0, np.zeros_like(h0), 5)
    assert rel(plan.step(H0, None, 5).cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("shape", [(130, 258), (258, 258)])
@pytest.mark.parametrize("cfg", list(range(1, 12)))
def test_gemm_tile_configs_agree(cfg, shape):
    """Every GEMM tile configuration gives the oracle's inversion (FP64).  At
    130 x 258 the folded K of 64 does not split four ways and 64 rows do not
    tile by 128, so some configurations fall back to the default; at 258 x 258
    every one applies."""
    ny, nx = shape
    rng = np.random.default_rng(3)
    orc = QgmOracle(dx=5e3, dy=6e3, dt=900.0, c=2.0, f=1e-4, time_scheme="rk2", shape=(ny, nx))
    plan = QgPlan(ny, nx, 5e3, 6e3, 900.0, 2.0, 1e-4, time_scheme="rk2", precision="float64",
                  gemm_config=cfg)
    q = 1e-5 * rng.standard_normal((ny, nx))
    hb = 0.01 * rng.standard_normal((ny, nx))
    qb = orc.h2pv(hb, hb)
    h = plan.pv2h(plan.field(q), plan.field(hb), plan.field(qb)).cpu().numpy()
    assert rel(h, orc.pv2h(q, hb, qb)) < 1e-12
    chosen = plan.autotune(reps=2)
    assert all(0 <= c < 11 for c in chosen)
    h2 = plan.pv2h(plan.field(q), plan.field(hb), plan.field(qb)).cpu().numpy()
    assert rel(h2, orc.pv2h(q, hb, qb)) < 1e-12


@pytest.mark.parametrize("cfg", list(range(1, 12)))
def test_gemm_tile_configs_f32(cfg):
    """Every tile configuration on the FP32 MFMA path (FP32 rounding over four
    sine products: 1e-5 relative)."""
    ny, nx = 258, 258
    rng = np.random.default_rng(4)
    orc = QgmOracle(dx=5e3, dy=6e3, dt=900.0, c=2.0, f=1e-4, time_scheme="rk2", shape=(ny, nx))
    plan = QgPlan(ny, nx, 5e3, 6e3, 900.0, 2.0, 1e-4, time_scheme="rk2", precision="float32",
                  gemm_config=cfg)
    q = 1e-5 * rng.standard_normal((ny, nx))
    hb = 0.01 * rng.standard_normal((ny, nx))
    qb = orc.h2pv(hb, hb)
    h = plan.pv2h(plan.field(q), plan.field(hb), plan.field(qb)).cpu().numpy()
    assert rel(h, orc.pv2h(q, hb, qb)) < 1e-5


@pytest.mark.parametrize("opts", [dict(), dict(land=True, beta=True, kdiff=True, bathy=True)])
def test_rhs_stage_f64(opts):
    """vd_qg_rhs: rhs(h, q) alone and fused with an Euler update, against the
    oracle's rhs (jqgm.py:321-375)."""
    orc, plan, h0, hb, rng = make_pair(36, 44, "Euler", "float64", **opts)
    q = orc.h2pv(h0, hb)
    ref = orc.rhs(*orc.h2uv(h0), q)
    H, Q = plan.field(h0), plan.field(q)
    r = plan.rhs(H, Q).cpu().numpy()
    assert rel(r, ref) < 1e-12
    e = plan.rhs(H, Q, base=Q, coef=900.0).cpu().numpy()
    assert rel(e, q + 900.0 * ref) < 1e-13



@pytest.mark.parametrize("opts", [dict(), dict(land=True, beta=True, kdiff=True, bathy=True)])
@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("float32", 2e-5)])
def test_rhs_adj_against_oracle(precision, tol, opts):
    """vd_qg_rhs_adj is the transpose of vd_qg_rhs: the oracle's hand vjp of
    rhs(h2uv(h), q) (its own dot-product test pins it), with an accumulator."""
    ny, nx = 66, 98
    orc, plan, h0, hb, rng = make_pair(ny, nx, "rk2", precision, **opts)
    q0 = orc.h2pv(h0, hb)
    lam = rng.standard_normal((ny, nx))
    acc = rng.standard_normal((ny, nx))
    coef = 450.0
    adh_ref, adq_ref = orc._k_adj(coef * lam, h0, q0)
    lq, lh = plan.rhs_adj(plan.field(h0), plan.field(q0), plan.field(lam), coef, plan.field(acc))
    torch.cuda.synchronize()
    assert rel(lq.cpu().numpy(), acc + adq_ref) < tol
    assert rel(lh.cpu().numpy(), adh_ref) < tol
    # in place accumulation (acc aliases lq)
    lq2 = plan.field(acc)
    plan.rhs_adj(plan.field(h0), plan.field(q0), plan.field(lam), coef, lq2, lq2)
    torch.cuda.synchronize()
    assert torch.equal(lq2, lq)


def test_rhs_adj_rejects_aliasing():
    from vardyn_amd._lib import VardynError
    orc, plan, h0, hb, rng = make_pair(40, 48, "rk2", "float64")
    h, q = plan.field(h0), plan.field(orc.h2pv(h0, hb))
    lam = plan.field(rng.standard_normal((40, 48)))
    with pytest.raises(VardynError):
        plan.rhs_adj(h, q, lam, 1.0, None, lam, plan.empty())
    with pytest.raises(VardynError):
        plan.rhs_adj(h, q, lam, 1.0, None, plan.empty(), lam)
