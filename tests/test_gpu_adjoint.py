"""GPU tests of the hand-written adjoint (lem_field_los_adjoint) behind
lyceanem.differentiable.FarField: against autograd of the float64 closed form
(dense reference gradient) and float64 central finite differences (h = 1e-6)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-3


def _small_c4(radius=0.5):
    from lyceanem import scenes

    return scenes.conformal_c4(n_az_el=16, n_z_el=8, n_el=19, n_az=37, n_side=32, n_ring=8,
                               radius=radius)


def _mask(native, sc):
    bits, _ = native.visibility_dense(sc["src"], sc["snk"], sc["tris"], any_hit=True)
    return bits, ~native.unpack_bits(bits, len(sc["snk"]))


def _db_mask_loss_torch(E, in_beam, sidelobe_db=-25.0, ref_power=None, eps=1e-30):
    """The dB mask loss as a composition of torch operations (any dtype and
    device, differentiated by autograd): the reference for lem_db_mask_loss."""
    P = (E.real ** 2 + E.imag ** 2).sum(dim=-1)
    peak = P.detach().max() if ref_power is None else ref_power
    db = 10.0 * torch.log10(P / peak + eps)
    err = torch.where(in_beam, db, torch.clamp(db - sidelobe_db, min=0.0))
    return (err ** 2).mean()


def _loss64(oracle, w, pos, sc, visible, in_beam, ref_power=None):
    E = oracle.field_los_torch64(w, pos, torch.from_numpy(sc["snk"].astype(np.float64)),
                                 torch.from_numpy(sc["snk_n"].astype(np.float64)),
                                 torch.from_numpy(visible), sc["alpha"], sc["beta"])
    return _db_mask_loss_torch(E, torch.from_numpy(in_beam), ref_power=ref_power)


@pytest.mark.parametrize("K,n_beam,ref_power", [(5000, 300, None), (5000, 300, 3.0),
                                                 (7, 2, None), (300000, 900, None)])
def test_db_mask_loss_matches_torch_reference(native, K, n_beam, ref_power):
    """lem_db_mask_loss (loss and dloss/dE in one pass) against autograd of the
    torch composition in float64 on the same complex64 field: a beam, side lobes
    on both sides of the -25 dB bar, a zero-power sink; from one workgroup
    (K = 7) to more sinks than the 1,024 workgroups cover in one pass."""
    from lyceanem.differentiable import db_mask_loss

    rng = np.random.default_rng(31 + K)
    amp = 10.0 ** (rng.uniform(-4.5, 0.0, K) / 2)
    E0 = (amp[:, None] * (rng.normal(size=(K, 3)) + 1j * rng.normal(size=(K, 3)))).astype(
        np.complex64)
    E0[min(17, K - 1)] = 0
    in_beam = np.zeros(K, bool)
    in_beam[:n_beam] = True
    dev = native.device()
    E = torch.from_numpy(E0).to(dev).requires_grad_(True)
    loss = db_mask_loss(E, torch.from_numpy(in_beam).to(dev), ref_power=ref_power)
    loss.backward()
    E64 = torch.from_numpy(E0.astype(np.complex128)).requires_grad_(True)
    l64 = _db_mask_loss_torch(E64, torch.from_numpy(in_beam), ref_power=ref_power)
    l64.backward()
    g, g64 = E.grad.cpu().numpy(), E64.grad.numpy()
    rl = abs(loss.item() - l64.item()) / abs(l64.item())
    rg = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    live = np.abs(g64).sum(-1) > 0
    print(f"dB mask loss: loss {rl:.2e}, gradient {rg:.2e} normwise, {int(live.sum())} of {K} "
          f"sinks with a gradient", flush=True)
    assert 0.1 < live.mean() < 0.95
    assert (g[~live] == 0).all()
    assert rl < 1e-5 and rg < 1e-5


@pytest.mark.parametrize("radius", [0.5, 3.0])
def test_adjoint_vs_dense_reference_gradient(native, oracle, radius):
    """A 0.5 m and a 3 m cylinder of 128 elements at 10 GHz."""
    from lyceanem.differentiable import far_field, db_mask_loss

    sc = _small_c4(radius)
    bits, visible = _mask(native, sc)
    assert 0.1 < visible.mean() < 0.9
    dev = native.device()
    w = torch.from_numpy(sc["w"]).to(dev).requires_grad_(True)
    pos = torch.from_numpy(sc["src"]).to(dev).requires_grad_(True)
    snk = torch.from_numpy(sc["snk"]).to(dev)
    snk_n = torch.from_numpy(sc["snk_n"]).to(dev)
    in_beam = torch.from_numpy(sc["in_beam"]).to(dev)
    E = far_field(w, pos, snk, snk_n, hit_bits=bits, beta=sc["beta"])
    loss = db_mask_loss(E, in_beam)
    loss.backward()
    w64 = torch.from_numpy(sc["w"].astype(np.complex128)).requires_grad_(True)
    p64 = torch.from_numpy(sc["src"].astype(np.float64)).requires_grad_(True)
    l64 = _loss64(oracle, w64, p64, sc, visible, sc["in_beam"])
    l64.backward()
    gw, gp = w.grad.cpu().numpy(), pos.grad.cpu().numpy()
    rw = np.linalg.norm(gw - w64.grad.numpy()) / np.linalg.norm(w64.grad.numpy())
    rp = np.linalg.norm(gp - p64.grad.numpy()) / np.linalg.norm(p64.grad.numpy())
    print(f"adjoint vs dense float64 autograd: weights {rw:.2e}, positions {rp:.2e}", flush=True)
    assert abs(loss.item() - l64.item()) < 1e-4 * abs(l64.item())
    assert rw < GRAD_RTOL and rp < GRAD_RTOL


def test_adjoint_vs_central_finite_differences(native, oracle):
    """64 random parameters (real / imaginary weight parts and radial offsets),
    float64 central differences with h = 1e-6."""
    from lyceanem.differentiable import far_field, db_mask_loss

    sc = _small_c4()
    bits, visible = _mask(native, sc)
    S = len(sc["src"])
    nhat = torch.from_numpy(sc["src_n"].astype(np.float64))
    p0 = torch.from_numpy(sc["src"].astype(np.float64))
    w0 = torch.from_numpy(sc["w"].astype(np.complex128))
    dev = native.device()
    a = torch.from_numpy(sc["w"][:, 2].copy()).to(dev).requires_grad_(True)
    delta = torch.zeros(S, dtype=torch.float32, device=dev, requires_grad=True)
    zhat = torch.tensor([0, 0, 1], dtype=torch.complex64, device=dev)
    pos = torch.from_numpy(sc["src"]).to(dev) + delta[:, None] * torch.from_numpy(
        sc["src_n"]).to(dev)
    E = far_field(a[:, None] * zhat, pos, torch.from_numpy(sc["snk"]).to(dev),
                  torch.from_numpy(sc["snk_n"]).to(dev), hit_bits=bits, beta=sc["beta"])
    # a fixed reference power: finite differences must see the same loss function
    ref = float((E.real ** 2 + E.imag ** 2).sum(-1).max().item())
    loss = db_mask_loss(E, torch.from_numpy(sc["in_beam"]).to(dev), ref_power=ref)
    loss.backward()
    ga = a.grad.cpu().numpy()
    gd = delta.grad.cpu().numpy()
    rng = np.random.default_rng(9)
    kinds = rng.integers(0, 3, 64)
    idx = rng.integers(0, S, 64)
    h = 1e-6

    def L(da=None, dd=None):
        w = w0.clone()
        p = p0.clone()
        if da is not None:
            w[da[0], 2] += da[1]
        if dd is not None:
            p[dd[0]] += dd[1] * nhat[dd[0]]
        return _loss64(oracle, w, p, sc, visible, sc["in_beam"], ref_power=ref).item()

    fd, ad = [], []
    for kd, i in zip(kinds, idx):
        if kd == 0:
            fd.append((L(da=(i, h)) - L(da=(i, -h))) / (2 * h))
            ad.append(ga[i].real)
        elif kd == 1:
            fd.append((L(da=(i, 1j * h)) - L(da=(i, -1j * h))) / (2 * h))
            ad.append(ga[i].imag)
        else:
            fd.append((L(dd=(i, h)) - L(dd=(i, -h))) / (2 * h))
            ad.append(gd[i])
    fd, ad = np.array(fd), np.array(ad)
    scale = np.abs(fd).max()
    err = np.abs(fd - ad) / scale
    print(f"finite differences: max |fd - adjoint| / max|fd| = {err.max():.2e}", flush=True)
    assert scale > 0 and err.max() < GRAD_RTOL


def test_adjoint_flags_match_autograd(native, oracle):
    """Sink weights and the sink-tangent step are transposed correctly."""
    rng = np.random.default_rng(3)
    S, K = 40, 300
    src = rng.uniform(-0.5, 0.5, (S, 3)).astype(np.float32)
    snk = (rng.normal(size=(K, 3)) * 50).astype(np.float32)
    nrm = rng.normal(size=(K, 3)).astype(np.float32)
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    wk = (rng.normal(size=(K, 3)) + 1j * rng.normal(size=(K, 3))).astype(np.complex64)
    g = (rng.normal(size=(K, 3)) + 1j * rng.normal(size=(K, 3))).astype(np.complex64)
    beta = 2 * np.pi / 0.2
    gw, gp = native.field_los_adjoint(src, w, snk, nrm, g, snk_w=wk, beta=beta,
                                      sink_tangent=True)
    w64 = torch.from_numpy(w.astype(np.complex128)).requires_grad_(True)
    p64 = torch.from_numpy(src.astype(np.float64)).requires_grad_(True)
    E = oracle.field_los_torch64(w64, p64, torch.from_numpy(snk.astype(np.float64)),
                                 torch.from_numpy(nrm.astype(np.float64)),
                                 torch.ones((S, K), dtype=torch.bool), 0.0, beta)
    E = E * torch.from_numpy(wk.astype(np.complex128))
    nh = torch.from_numpy(nrm.astype(np.float64))
    nh = nh / nh.norm(dim=1, keepdim=True)
    E = E - nh * (E * nh).sum(-1, keepdim=True)
    E.backward(torch.from_numpy(g.astype(np.complex128)))
    rw = np.linalg.norm(gw.cpu().numpy() - w64.grad.numpy()) / np.linalg.norm(w64.grad.numpy())
    rp = np.linalg.norm(gp.cpu().numpy() - p64.grad.numpy()) / np.linalg.norm(p64.grad.numpy())
    assert rw < GRAD_RTOL and rp < GRAD_RTOL
    # lanes over the sources in input order: the same per-source sums
    gw2, gp2 = native.field_los_adjoint(src, w, snk, nrm, g, snk_w=wk, beta=beta,
                                        sink_tangent=True, input_order=True)
    assert np.array_equal(gw.cpu().numpy(), gw2.cpu().numpy())
    assert np.array_equal(gp.cpu().numpy(), gp2.cpu().numpy())


def test_far_field_plan_is_bit_identical(native):
    """FarFieldPlan runs the forward over a spatial sink order with the mask's
    columns permuted, and the adjoint over a mask-driven source order with the
    mask's rows permuted: each sink's sum is over the same sources, and each
    source's over the same sinks, in the same order, so field and gradients
    equal the unplanned ones bit for bit."""
    from lyceanem.differentiable import FarFieldPlan, far_field, db_mask_loss

    sc = _small_c4()
    bits, visible = _mask(native, sc)
    dev = native.device()
    snk = torch.from_numpy(sc["snk"]).to(dev)
    snk_n = torch.from_numpy(sc["snk_n"]).to(dev)
    in_beam = torch.from_numpy(sc["in_beam"]).to(dev)
    plan = FarFieldPlan(snk, snk_n, bits, center=sc["src"].mean(0))
    assert not torch.equal(plan.order.cpu(), torch.arange(len(snk)))
    # the adjoint over the mask-driven source order (forced) and the automatic choice
    plan_m = FarFieldPlan(snk, snk_n, bits, center=sc["src"].mean(0), adjoint_order="mask")
    assert plan_m.src_order is not None
    assert not torch.equal(plan_m.src_order.cpu(), torch.arange(len(sc["src"])))
    plan_a = FarFieldPlan(snk, snk_n, bits, center=sc["src"].mean(0),
                          sources=torch.from_numpy(sc["src"]).to(dev))
    print(f"plan with sources: forward {plan_a.sink_order} {plan_a.forward_times}, adjoint "
          f"{plan_a.adjoint_order} {plan_a.adjoint_times}", flush=True)
    plan_z = FarFieldPlan(snk, snk_n, bits, center=sc["src"].mean(0), sink_order="azimuth",
                          adjoint_order="input")
    assert plan_z.sink_order == "azimuth" and plan_z.src_order is not None
    out = []
    for p in (None, plan, plan_m, plan_a, plan_z):
        w = torch.from_numpy(sc["w"]).to(dev).requires_grad_(True)
        pos = torch.from_numpy(sc["src"]).to(dev).requires_grad_(True)
        E = far_field(w, pos, snk, snk_n, hit_bits=bits, beta=sc["beta"], plan=p)
        loss = db_mask_loss(E, in_beam)
        loss.backward()
        out.append((E.detach().cpu(), w.grad.cpu(), pos.grad.cpu()))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert torch.equal(a, b)


def test_gradient_step_replays_in_a_graph(native):
    """The whole gradient step -- planned forward, loss, hand-written adjoint --
    captured in a HIP graph replays bit for bit like the eager step, replay
    after replay, with eager work in between (no host synchronisation or
    memset node inside the step)."""
    from lyceanem.differentiable import FarFieldPlan, far_field, db_mask_loss

    sc = _small_c4()
    bits, _ = _mask(native, sc)
    dev = native.device()
    snk = torch.from_numpy(sc["snk"]).to(dev)
    snk_n = torch.from_numpy(sc["snk_n"]).to(dev)
    in_beam = torch.from_numpy(sc["in_beam"]).to(dev)
    plan = FarFieldPlan(snk, snk_n, bits, adjoint_order="mask")
    w0 = torch.from_numpy(sc["w"]).to(dev)
    p0 = torch.from_numpy(sc["src"]).to(dev)

    def step(w, pos):
        E = far_field(w, pos, snk, snk_n, hit_bits=bits, beta=sc["beta"], plan=plan)
        loss = db_mask_loss(E, in_beam)
        loss.backward()
        return loss.detach()

    def eager(scale):
        w = (w0 * scale).requires_grad_(True)
        pos = p0.clone().requires_grad_(True)
        loss = step(w, pos)
        return loss, w.grad, pos.grad

    wg = w0.clone().requires_grad_(True)
    pg = p0.clone().requires_grad_(True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            wg.grad = None
            pg.grad = None
            step(wg, pg)
    torch.cuda.current_stream().wait_stream(side)
    wg.grad = None
    pg.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        g_loss = step(wg, pg)
    for i, scale in enumerate((1.0, 0.5, 1.0, 2.0)):
        ref = eager(scale)
        torch.cuda.synchronize()
        with torch.no_grad():
            wg.copy_(w0 * scale)
        graph.replay()
        torch.cuda.synchronize()
        got = (g_loss, wg.grad, pg.grad)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), f"replay {i}"
