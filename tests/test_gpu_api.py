"""The reference-style surfaces on the GPU: the ``Qgm`` drop-in (jqgm.py),
``Model_qg1l`` with ``State`` (mod.py ``M`` protocol), ``Obsop_interp_l3``
(obsop.py ``misfit`` / ``adj``), the generic ``Variational.cost_and_grad``
segment loop (tools_4Dvar.py:252-400) and the identity tests (mod.py:5327-5446,
tools_4Dvar.py:402-413).

Bars: FP64 results within 1e-10 relative of the FP64 oracle; the generic
``Variational`` loop and the preallocated ``FourDVar`` engine agree to 1e-12
(same kernels, different host-side summation order of J); the adjoint test
ratio within 1e-11 of 1.
"""
import numpy as np
import pytest
import torch

from oracle.fourdvar import ModelOracle
from oracle.obsop import ObsopOracle
from oracle.qg import QgmOracle
from vardyn_amd import synthetic as syn
from vardyn_amd.basis import Basis_ic, Basis_offset
from vardyn_amd.jqgm import Qgm
from vardyn_amd.mod import Model_qg1l, adjoint_test, tangent_test
from vardyn_amd.problem import build_problem
from vardyn_amd.state import State
from vardyn_amd.tools_4Dvar import Cov, Variational, grad_test

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN pattern differs"
    m = ~np.isnan(b)
    return np.linalg.norm(a[m] - b[m]) / max(np.linalg.norm(b[m]), 1e-300)


def reference_style_args(ny=40, nx=46, land=True):
    """Arguments as Model_qg1l_jax passes them to Qgm (mod.py:697-718): 2-D
    DX/DY, an SSH field with NaN on land, c and f as 2-D fields."""
    DX = np.full((ny, nx), 9e3)
    DY = np.full((ny, nx), 11e3)
    ssh = np.zeros((ny, nx))
    if land:
        ssh[ny // 2, nx // 3] = np.nan
        ssh[ny // 2, nx // 3 + 1] = np.nan
        ssh[4, nx - 7] = np.nan
    c = np.full((ny, nx), 2.5)
    f = 1e-4 + 2e-11 * 11e3 * (np.arange(ny)[:, None] - ny / 2) * np.ones((1, nx))
    return dict(dx=DX, dy=DY, dt=1200.0, SSH=ssh, c=c, f=f, Kdiffus=30.0, time_scheme="rk2")


def test_qgm_dropin_numpy_in_numpy_out():
    kw = reference_style_args()
    qg = Qgm(precision="float64", upwind=3, g=9.81, compile=True, Wbc=None, advect_pv=True,
             **kw)
    orc = QgmOracle(**kw)
    rng = np.random.default_rng(3)
    ny, nx = kw["SSH"].shape
    h0 = 0.1 * rng.standard_normal((ny, nx))
    h0[np.isnan(kw["SSH"])] = np.nan
    hb = 0.01 * rng.standard_normal((ny, nx))
    h1 = qg.step_jit(h0, hb, nstep=3)
    assert isinstance(h1, np.ndarray)
    ref = orc.step(h0, hb, 3)
    assert rel(h1, ref) < 1e-10
    assert np.all(np.isnan(h1[qg.ind0]))

    dh = 0.1 * rng.standard_normal((ny, nx))
    tl = qg.step_tgl_jit(dh, h0, hb, nstep=3)
    tl_ref, _ = orc.step_tgl(dh, h0, hb, 3)
    assert rel(tl, tl_ref) < 1e-10

    ay = rng.standard_normal((ny, nx))
    ad = qg.step_adj_jit(ay, h0, hb, nstep=3)
    ad_ref = orc.step_adj(ay, h0, hb, 3)
    assert rel(ad, ad_ref) < 1e-10


def test_qgm_dropin_torch_stays_on_device():
    kw = reference_style_args(land=False)
    qg = Qgm(precision="float64", **kw)
    h0 = torch.zeros(kw["SSH"].shape, dtype=torch.float64, device="cuda")
    h1 = qg.step_jit(h0, torch.zeros_like(h0), nstep=2)
    assert isinstance(h1, torch.Tensor) and h1.is_cuda
    # a state at rest with zero boundary stays at rest
    assert torch.count_nonzero(torch.nan_to_num(h1)).item() == 0


def test_qgm_dropin_float32_default(monkeypatch):
    monkeypatch.delenv("VARDYN_FLOAT64", raising=False)
    kw = reference_style_args(land=False)
    qg = Qgm(**kw)
    assert qg.dtype == "float32"
    rng = np.random.default_rng(4)
    h0 = 0.1 * rng.standard_normal(kw["SSH"].shape)
    hb = np.zeros_like(h0)
    h1 = qg.step_jit(h0, hb, nstep=2)
    assert h1.dtype == np.float32
    assert rel(h1, QgmOracle(**kw).step(h0, hb, 2)) < 1e-4


def make_model(ny=36, nx=42, land=False, precision="float64"):
    kw = reference_style_args(ny, nx, land=land)
    st = State(ny, nx, dtype=torch.float64, DX=kw["dx"], DY=kw["dy"])
    M = Model_qg1l(st, kw["dt"], c=kw["c"], f=kw["f"], time_scheme="rk2", Kdiffus=kw["Kdiffus"],
                   SSH=kw["SSH"], precision=precision, final_date=40 * kw["dt"])
    return M, st, kw


def test_model_step_with_flux_control_vs_oracle():
    M, st, kw = make_model(land=True)
    orc = ModelOracle(QgmOracle(**kw))
    rng = np.random.default_rng(5)
    h0 = 0.1 * rng.standard_normal((st.ny, st.nx))
    h0[np.isnan(kw["SSH"])] = np.nan
    flux = 0.02 * rng.standard_normal((st.ny, st.nx))
    st.setvar(h0, "ssh")
    st.params["ssh"] = torch.as_tensor(flux, device=st.device)
    M.step(st, nstep=4, t=0)
    ref = orc.step(h0, flux, 4)
    assert rel(st.getvar("ssh"), ref) < 1e-10


def test_model_tangent_and_adjoint_identities():
    M, st, _ = make_model()
    ratios = tangent_test(M, st, nstep=3, ampl=0.1, lambdas=range(6))
    # first-order Taylor remainder: the ratio falls by ~10x per decade of lambda
    assert ratios[-1] < 1e-4 and ratios[-1] < ratios[0] * 1e-3
    r = adjoint_test(M, st, nstep=3, ampl=0.1)
    assert abs(r - 1) < 1e-11


def test_model_boundary_fields_nearest_time():
    M, st, kw = make_model()
    rng = np.random.default_rng(6)
    hbs = [0.01 * rng.standard_normal((st.ny, st.nx)) for _ in range(2)]
    M.set_bc([0.0, 3600.0 * 6], {"SSH": hbs})
    h0 = 0.1 * rng.standard_normal((st.ny, st.nx))
    st.setvar(h0, "ssh")
    M.step(st, nstep=2, t=3600.0 * 5)   # closest boundary time is 6 h
    ref = QgmOracle(**kw).step(h0, hbs[1], 2)
    assert rel(st.getvar("ssh"), ref) < 1e-10


def small_case():
    return syn.QgCase("api", 48, 56, 10e3, 2.7, 1200.0, 36, n_tracks=5, track_km=7.0,
                      checkpoint_steps=9, time_scheme="rk2", seed=13)


def test_obsop_misfit_and_adjoint_vs_oracle():
    pb = build_problem(small_case(), precision="float64", with_engine=False)
    c = pb.case
    H = pb.H
    Ho = ObsopOracle(pb.grid.lon, pb.grid.lat, pb.grid.DX, pb.grid.DY, pb.T, c.dt, pb.dict_obs,
                     Npix=4)
    rng = np.random.default_rng(8)
    x = 0.1 * rng.standard_normal((c.ny, c.nx))
    x[5, 7] = np.nan   # a NaN pixel drops the observations that touch it
    st = State(c.ny, c.nx, dtype=torch.float64)
    st.setvar(x, "ssh")
    for t in H.t_obs:
        m = H.misfit(t, st)
        mo = Ho.misfit(t, x)
        # the oracle pads every time's table to the largest count (obsop.py:583-604);
        # padded entries (y = 0, sigma = 1e7, weight 0) give a zero misfit
        n = m.numel()
        assert np.allclose(m.cpu().numpy(), mo[:n], rtol=1e-13, atol=1e-13)
        assert not np.any(mo[n:])
        ad = st.copy(free=True)
        H.adj(t, ad, st, m)
        ado = Ho.adj(t, x, mo)
        assert rel(ad.getvar("ssh"), ado) < 1e-13


def generic_and_engine(control):
    pb = build_problem(small_case(), precision="float64", control=control)
    c = pb.case
    st = State(c.ny, c.nx, dtype=torch.float64, DX=pb.grid.DX, DY=pb.grid.DY)
    M = Model_qg1l(st, c.dt, c=c.c, f=c.f0, time_scheme=c.time_scheme, precision="float64",
                   final_date=c.n_steps * c.dt)
    if control == "ic":
        basis, sig = Basis_ic(st, c.sigma_b), c.sigma_b
    else:
        basis, sig = Basis_offset(st, 0.01), 0.01
    var = Variational(M=M, H=pb.H, State=st, B=Cov(sig), Basis=basis,
                      Xb=np.zeros(basis.nbasis), checkpoints=pb.checkpoints,
                      dtbasis=c.checkpoint_steps)
    return pb, var


@pytest.mark.parametrize("control", ["ic", "offset"])
def test_variational_generic_equals_engine(control):
    pb, var = generic_and_engine(control)
    rng = np.random.default_rng(9)
    X = (0.05 * rng.standard_normal(pb.engine.nctrl)) if control == "ic" else np.array([0.02])
    J, G = var.cost_and_grad(X)
    Je, Ge = pb.engine.cost_and_grad(X)
    assert J == pytest.approx(Je, rel=1e-12)
    assert rel(G, Ge) < 1e-12


def test_grad_test_on_engine():
    pb = build_problem(small_case(), precision="float64", control="ic")
    eng = pb.engine
    X = 0.05 * np.random.default_rng(10).standard_normal(eng.nctrl)
    r = grad_test(lambda x: eng.cost_and_grad(x)[0], lambda x: eng.cost_and_grad(x)[1], X, n=7)
    # 1 - (J(X+lh) - J(X)) / (l G.h) = O(l): a tenfold drop per decade of l
    # (FP64 cancellation stays far below at these l)
    assert r[-1] < 1e-5
    drop = r[-2:] / r[-3:-1]   # the last two decades, past the pre-asymptotic range
    assert np.all(drop < 0.2) and np.all(drop > 0.05)


def test_autograd_backward_is_the_adjoint():
    """QgStep.backward = the hand adjoint: h0.grad of <h1, w> equals the
    oracle's step_adj(w) (jqgm.py:565-569)."""
    from vardyn_amd.autograd import qg_step
    from vardyn_amd.qg import QgPlan
    ny, nx, n = 40, 46, 4
    orc = QgmOracle(dx=9e3, dy=11e3, dt=1200.0, c=2.5, f=1e-4, time_scheme="rk3", shape=(ny, nx))
    plan = QgPlan(ny, nx, 9e3, 11e3, 1200.0, 2.5, 1e-4, time_scheme="rk3", precision="float64")
    rng = np.random.default_rng(11)
    h0n, hbn, wn = (0.1 * rng.standard_normal((ny, nx)), 0.01 * rng.standard_normal((ny, nx)),
                    rng.standard_normal((ny, nx)))
    h0 = plan.field(h0n).requires_grad_(True)
    hb, w = plan.field(hbn), plan.field(wn)
    h1 = qg_step(h0, hb, plan, n)
    assert rel(h1.detach().cpu().numpy(), orc.step(h0n, hbn, n)) < 1e-10
    (h1 * w).sum().backward()
    assert rel(h0.grad.cpu().numpy(), orc.step_adj(wn, h0n, hbn, n)) < 1e-10


def test_autograd_jvp_is_the_tangent_linear():
    from vardyn_amd.autograd import qg_step
    from vardyn_amd.qg import QgPlan
    ny, nx, n = 40, 46, 4
    orc = QgmOracle(dx=9e3, dy=11e3, dt=1200.0, c=2.5, f=1e-4, time_scheme="rk2", shape=(ny, nx))
    plan = QgPlan(ny, nx, 9e3, 11e3, 1200.0, 2.5, 1e-4, time_scheme="rk2", precision="float64")
    rng = np.random.default_rng(12)
    h0n, dhn = 0.1 * rng.standard_normal((ny, nx)), 0.1 * rng.standard_normal((ny, nx))
    import torch.autograd.forward_ad as fwAD
    with fwAD.dual_level():
        y = qg_step(fwAD.make_dual(plan.field(h0n), plan.field(dhn)), None, plan, n)
        h1, dh1 = fwAD.unpack_dual(y)
    tl_ref, h1_ref = orc.step_tgl(dhn, h0n, np.zeros((ny, nx)), n)
    assert rel(h1.cpu().numpy(), orc.step(h0n, np.zeros((ny, nx)), n)) < 1e-10
    assert rel(dh1.cpu().numpy(), tl_ref) < 1e-10


def test_autograd_rejects_boundary_gradient():
    from vardyn_amd._lib import VardynError
    from vardyn_amd.autograd import qg_step
    from vardyn_amd.qg import QgPlan
    plan = QgPlan(24, 24, 1e4, 1e4, 900.0, 2.0, 1e-4, precision="float64")
    h0 = torch.zeros((24, 24), dtype=torch.float64, device="cuda", requires_grad=True)
    hb = torch.zeros_like(h0, requires_grad=True)
    with pytest.raises(VardynError):
        qg_step(h0, hb, plan, 1)
