"""GPU checks of the FP32-storage adjoint (snapx_adjoint_f32_*).

FP32 tapes and adjoint fields, FP64 arithmetic in registers and FP64 column
solves.  What pins it is the FP64 adjoint (itself pinned by central
differences, gradcheck and libtorch autograd through the torch oracle):

* one step: the FP32 adjoint of a random weighting against the FP64 adjoint
  of the same weighting at the same state, per variable relative to the
  variable's largest entry;
* record + sweep gives the bits of snapx_adjoint_f32_rk3_step, and the
  recorded step output follows the FP64 step to FP32 rounding;
* seven steps through ``integrate`` with taped segments (fused FP32 forward
  steps between the checkpoints): the gradient against the FP64 one.
"""
import numpy as np
import pytest
import torch

from snapy_linfel_amd import capi
from snapy_linfel_amd import adjoint as A
from test_gpu_adjoint import _case, _scales, _step
from test_gpu_plain import cfg_of, to_dev

pytestmark = pytest.mark.gpu

CASES = ["straka1", "straka9", "box_lmars_roe", "box_hllc", "straka_mid1", "straka_tall1",
         "sph3d_hllc1", "sph2d_roe9", "sph2dtall_hllc1"]


def _lam(u0, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(u0.shape, generator=g, dtype=torch.float64).to(u0.device) / _scales(u0)


def _rel(got, ref):
    """max |got - ref| per variable over the variable's largest |ref|"""
    sc = ref.abs().amax(dim=(1, 2, 3))
    sc = torch.maximum(sc, sc.max() * 1e-12)
    return ((got - ref).abs().amax(dim=(1, 2, 3)) / sc).cpu().numpy()


@pytest.mark.parametrize("case", CASES)
def test_f32_adjoint_close_to_fp64(case):
    """dL/du_in of one step from the FP32-storage adjoint within 1e-4 of the
    FP64 adjoint (largest entry per variable).  Measured: 1e-7 to 3e-6 on the
    Straka and HLLC cases, 4e-5 for the momenta of the LMARS / Roe box, whose
    momenta are small next to rho c: the FP32 rounding of the total state
    (about 1e-7 of rho c) is then 1e-5 of the momentum field itself."""
    oc, u_np, dt = _case(case)
    cfg = cfg_of(oc)
    u64 = to_dev(u_np)
    lam = _lam(u64, 7)
    l64 = lam.clone()
    capi.adjoint_rk3_step(cfg, dt, u64, l64)
    l32 = lam.float()
    capi.adjoint_rk3_step(cfg, dt, u64.float(), l32)
    err = _rel(l32.double(), l64)
    print(case, "FP32 adjoint vs FP64, per variable", err)
    assert np.isfinite(err).all() and (err < 1e-4).all(), err


@pytest.mark.parametrize("case", ["straka9", "box_lmars_roe"])
def test_f32_record_sweep_bits_and_step_output(case):
    oc, u_np, dt = _case(case)
    cfg = cfg_of(oc)
    u0 = to_dev(u_np, torch.float32)
    lam = _lam(to_dev(u_np), 8).float()
    nb = capi.adjoint_tape_bytes(cfg, torch.float32)
    assert 0 < nb < capi.adjoint_tape_bytes(cfg)
    tape = torch.empty(nb, dtype=torch.uint8, device=u0.device)
    work = capi.adjoint_work(cfg, u0.device, torch.float32)
    out = torch.empty_like(u0)
    l1, l2 = lam.clone(), lam.clone()
    capi.adjoint_record(cfg, dt, u0, out, tape, work)
    capi.adjoint_sweep(cfg, dt, tape, l1, work)
    capi.adjoint_rk3_step(cfg, dt, u0, l2)
    torch.cuda.synchronize()
    assert torch.equal(l1, l2)
    # the recorded FP32 step against the FP64 step (the same total-state
    # rounding: 2e-5 of the box's momenta)
    ref = _step(cfg, dt, to_dev(u_np))
    rel = ((out.double() - ref).abs() / _scales(ref)).amax()
    assert rel < 1e-4, float(rel)


def test_f32_mixed_dtypes_rejected():
    oc, u_np, dt = _case("straka1")
    cfg = cfg_of(oc)
    with pytest.raises(capi.SnapxError):
        capi.adjoint_rk3_step(cfg, dt, to_dev(u_np, torch.float32), to_dev(u_np))


@pytest.mark.parametrize("case", ["straka1", "box_lmars_roe"])
def test_f32_taped_integrate_close_to_fp64(case):
    """Seven steps, checkpoints every three, taped segments: the FP32
    gradient within 1e-4 of the FP64 one (largest entry per variable)."""
    oc, u_np, dt = _case(case)
    cfg = cfg_of(oc)
    u64 = to_dev(u_np)
    lam = _lam(u64, 9)
    grads = []
    for x0, lw in ((u64, lam), (u64.float(), lam.float())):
        x = x0.clone().requires_grad_(True)
        y = A.integrate(x, cfg, dt, 7, checkpoint_every=3)
        (gr,) = torch.autograd.grad((y * lw).sum(), x)
        grads.append(gr.double())
    err = _rel(grads[1], grads[0])
    print(case, "FP32 taped gradient vs FP64, per variable", err)
    assert np.isfinite(err).all() and (err < 1e-4).all(), err


@pytest.mark.parametrize("scheme", [1, 9])
def test_batch_gradient_matches_members(scheme):
    """integrate_batch over three members of different shapes (two 2-D
    Straka blocks and a 3-D box): each member's slice of the gradient equals
    that member's own FP32 taped gradient to 1e-5 (largest entry per
    variable); the batch's Jacobian is block-diagonal."""
    from test_gpu_plain import _box3d, _straka_small
    ocs = [_straka_small(scheme, nx2=24, nx1=16, wind=10.0)[0::2],
           _straka_small(scheme, nx2=16, nx1=20, wind=5.0)[0::2],
           _box3d(n=(10, 8, 6), scheme=scheme)[0::2]]
    cfgs = [cfg_of(oc) for oc, _ in ocs]
    dt = 0.2
    b = capi.Batch(cfgs, dt)
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, u0 in ocs])
    g = torch.Generator(device="cpu").manual_seed(11)
    lam = torch.randn(u.shape, generator=g, dtype=torch.float32).to(u.device)
    for m in range(b.n):
        b.member(lam, m).div_(_scales(b.member(u, m).double()).float())
    x = u.clone().requires_grad_(True)
    y = A.integrate_batch(x, b, dt, 4, checkpoint_every=2)
    (gb,) = torch.autograd.grad((y * lam).sum(), x)
    for m, (oc, u0) in enumerate(ocs):
        xm = to_dev(u0, torch.float32).requires_grad_(True)
        ym = A.integrate(xm, cfgs[m], dt, 4, checkpoint_every=2)
        (gm,) = torch.autograd.grad((ym * b.member(lam, m)).sum(), xm)
        err = _rel(b.member(gb, m).double(), gm.double())
        print("member", m, "batch vs single-member gradient", err)
        assert np.isfinite(err).all() and (err < 1e-5).all(), (m, err)
