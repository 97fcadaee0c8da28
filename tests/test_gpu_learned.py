"""GPU checks of the learned preconditioner's own kernels (BF16 MFMA
convolution forward, data and weight gradients, pooling, Adam, the A-norm
loss) against PyTorch CPU autograd in FP64 on the same bf16-rounded inputs,
and of training and flexible CG end to end.

Tolerances: bf16 activations carry 8 mantissa bits, so outputs are compared
at 2e-2 of their magnitude and gradients by cosine similarity >= 0.995."""
import ctypes as C

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import oracle as O
from pyom2_amd import _lib, basins, learned
from pyom2_amd.learned import LAYERS, PrecondNet
from pyom2_amd.streamfunction import SolveOptions, StreamfunctionSolver, islands_trace

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bf(x):
    """Round to bfloat16 (returns float64 CPU tensor of the rounded values)."""
    return torch.as_tensor(x).to(torch.bfloat16).to(torch.float64)


def p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def cos(a, b):
    a, b = a.flatten().double(), b.flatten().double()
    return float((a @ b) / (a.norm() * b.norm() + 1e-300))


def test_mfma_lane_layout():
    L = _lib.lib()
    g = torch.Generator().manual_seed(0)
    A = torch.randn(16, 32, generator=g).to(torch.bfloat16)
    B = torch.randn(32, 16, generator=g).to(torch.bfloat16)
    D = torch.zeros(16, 16, device=DEV)
    Ad, Bd = A.to(DEV).contiguous(), B.to(DEV).contiguous()
    assert L.psinn_mfma_probe(p(Ad), p(Bd), p(D), None) == 0
    torch.cuda.synchronize()
    ref = A.double() @ B.double()
    np.testing.assert_allclose(D.cpu().double().numpy(), ref.numpy(), rtol=1e-5, atol=1e-5)


def _conv_case(Bn, H, W, c1, up1, c2, cout, relu, seed=1):
    g = torch.Generator().manual_seed(seed)
    h1, w1 = (H // 2, W // 2) if up1 else (H, W)
    in1 = torch.randn(Bn, h1, w1, c1, generator=g).to(torch.bfloat16)
    in2 = torch.randn(Bn, H, W, c2, generator=g).to(torch.bfloat16) if c2 else None
    cin = c1 + c2
    cpad = (cout + 15) // 16 * 16
    w = torch.randn(9, cin, cpad, generator=g) * (1.0 / np.sqrt(9 * cin))
    w[:, :, cout:] = 0
    b = torch.randn(cpad, generator=g) * 0.1
    return in1, in2, w, b, cpad


def _torch_conv(in1, in2, w, b, up1, relu, cout):
    x1 = in1.double().permute(0, 3, 1, 2)
    if up1:
        x1 = x1.repeat_interleave(2, 2).repeat_interleave(2, 3)
    x = torch.cat([x1, in2.double().permute(0, 3, 1, 2)], 1) if in2 is not None else x1
    wt = bf(w).permute(2, 1, 0).reshape(w.shape[2], w.shape[1], 3, 3)[:cout]
    y = F.conv2d(x, wt, b[:cout].double(), padding=1)
    return torch.relu(y) if relu else y


@pytest.mark.parametrize("case", [(2, 16, 24, 32, True, 16, 48, True),
                                  (1, 8, 40, 16, False, 0, 32, True),
                                  (2, 12, 16, 64, True, 32, 32, False),
                                  # several 4 x 64 tiles with ragged edges; 192 input
                                  # channels (the data gradient has 12 output blocks)
                                  (1, 10, 136, 128, True, 64, 64, True),
                                  (1, 9, 70, 16, False, 0, 16, True)])
def test_conv3x3_forward_and_gradients(case):
    Bn, H, W, c1, up1, c2, cout, relu = case
    L = _lib.lib()
    in1, in2, w, b, cpad = _conv_case(Bn, H, W, c1, up1, c2, cout, relu)
    cin = c1 + c2
    d1, d2 = in1.to(DEV).contiguous(), (in2.to(DEV).contiguous() if in2 is not None else None)
    wd, bd = w.to(DEV).contiguous(), b.to(DEV).contiguous()
    pf = torch.empty(learned.pack_elems(cin // 16, cpad // 16), dtype=torch.int16, device=DEV)
    pt = torch.empty_like(pf)
    assert L.psinn_pack_weights(p(wd), cin, cpad, 0, cin // 16, cpad // 16, p(pf), None) == 0
    assert L.psinn_pack_weights(p(wd), cin, cpad, 1, cpad // 16, cin // 16, p(pt), None) == 0
    out = torch.empty(Bn, H, W, cout, dtype=torch.float32, device=DEV)
    assert L.psinn_conv3x3(p(d1), c1, int(up1), p(d2), c2, Bn, H, W, p(pf), cpad, p(bd), int(relu),
                           None, p(out), cout, None) == 0
    torch.cuda.synchronize()
    # forward
    x1 = in1.double().requires_grad_(True)
    x2 = in2.double().requires_grad_(True) if in2 is not None else None
    wt = bf(w).clone().requires_grad_(True)
    bt = b.double().clone().requires_grad_(True)
    X1 = x1.permute(0, 3, 1, 2)
    if up1:
        X1 = X1.repeat_interleave(2, 2).repeat_interleave(2, 3)
    X = torch.cat([X1, x2.permute(0, 3, 1, 2)], 1) if x2 is not None else X1
    Wt = wt.permute(2, 1, 0).reshape(cpad, cin, 3, 3)
    Y = F.conv2d(X, Wt, bt, padding=1)
    if relu:
        Y = torch.relu(Y)
    ref = Y[:, :cout].permute(0, 2, 3, 1)
    got = out.cpu().double()
    scale = ref.abs().max().item()
    assert (got - ref.detach()).abs().max().item() <= 2e-2 * scale
    # gradients for an upstream gradient G on the pre-activation (bf16)
    g = torch.Generator().manual_seed(7)
    G = torch.randn(Bn, H, W, cpad, generator=g).to(torch.bfloat16)
    G[..., cout:] = 0
    Gd = G.to(DEV).contiguous()
    Ypre = F.conv2d(X, Wt, bt, padding=1)
    (Ypre * G.double().permute(0, 3, 1, 2)).sum().backward()
    dW = torch.zeros(9, cin, cpad, device=DEV)
    db = torch.zeros(cpad, device=DEV)
    assert L.psinn_conv3x3_wgrad(p(d1), c1, int(up1), p(d2), c2, Bn, H, W, p(Gd), cpad, p(dW), p(db),
                                 None) == 0
    gin = torch.empty(Bn, H, W, cin, dtype=torch.float32, device=DEV)
    assert L.psinn_conv3x3(p(Gd), cpad, 0, None, 0, Bn, H, W, p(pt), cin, None, 0, None, p(gin), cin,
                           None) == 0
    torch.cuda.synchronize()
    assert cos(dW.cpu(), wt.grad) >= 0.9999
    assert (dW.cpu().double() - wt.grad).abs().max() <= 1e-3 * wt.grad.abs().max()
    assert cos(db.cpu(), bt.grad) >= 0.9999
    # data gradient at the conv resolution, per input part
    g1 = gin.cpu().double()
    Xl = X.detach().clone().requires_grad_(True)
    (F.conv2d(Xl, Wt.detach(), bt.detach(), padding=1) * G.double().permute(0, 3, 1, 2)).sum().backward()
    ref_g = Xl.grad.permute(0, 2, 3, 1)
    assert (g1 - ref_g).abs().max() <= 1e-3 * ref_g.abs().max()


def _torch_net(net: PrecondNet, x16):
    """FP64 CPU reimplementation of the network on the same bf16 weights."""
    Wt, Bt = {}, {}
    for ly in LAYERS:
        w = bf(net.w[ly.name].cpu()).clone().requires_grad_(True)
        b = net.b[ly.name].cpu().double().clone().requires_grad_(True)
        Wt[ly.name], Bt[ly.name] = w, b

    def conv(name, x, relu=True):
        ly = {l.name: l for l in LAYERS}[name]
        w = Wt[name].permute(2, 1, 0).reshape(ly.cout_pad, ly.cin, 3, 3)[:ly.cout]
        y = F.conv2d(x, w, Bt[name][:ly.cout], padding=1)
        return torch.relu(y) if relu else y

    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3)
    x = x16.double().permute(0, 3, 1, 2)
    a0 = conv("E0", x)
    a1 = conv("E1", F.avg_pool2d(a0, 2))
    a2 = conv("E2", F.avg_pool2d(a1, 2))
    a3 = conv("E3", F.avg_pool2d(a2, 2))
    d2 = conv("D2", torch.cat([up(a3), a2], 1))
    d1 = conv("D1", torch.cat([up(d2), a1], 1))
    d0 = conv("D0", torch.cat([up(d1), a0], 1))
    z = conv("OUT", d0, relu=False)[:, 0]
    return z, Wt, Bt


def test_network_forward_backward_matches_autograd():
    net = PrecondNet(device=DEV, seed=3)
    Bn, H, W = 2, 32, 48
    g = torch.Generator().manual_seed(5)
    x = torch.zeros(Bn, H, W, 16)
    x[..., :3] = torch.randn(Bn, H, W, 3, generator=g)
    x16 = x.to(torch.bfloat16)
    z = net.forward(x16.to(DEV).contiguous(), Bn, H, W)
    torch.cuda.synchronize()
    zr, Wt, Bt = _torch_net(net, x16)
    zg = z.cpu().double()
    assert cos(zg, zr.detach()) >= 0.995
    G = torch.randn(Bn, H, W, generator=g)
    net.zero_grad()
    net.backward(G.to(DEV).contiguous())
    torch.cuda.synchronize()
    (zr * G.double()).sum().backward()
    for ly in LAYERS:
        got, ref = net.gw[ly.name].cpu(), Wt[ly.name].grad
        c = cos(got, ref)
        assert c >= 0.99, (ly.name, c)
        assert cos(net.gb[ly.name].cpu(), Bt[ly.name].grad) >= 0.99, ly.name


def test_adam_matches_reference_formula():
    L = _lib.lib()
    g = torch.Generator().manual_seed(2)
    w = torch.randn(1000, generator=g)
    grads = [torch.randn(1000, generator=g) for _ in range(3)]
    wd, m, v = w.to(DEV).clone(), torch.zeros(1000, device=DEV), torch.zeros(1000, device=DEV)
    ref = w.double().clone()
    mr, vr = torch.zeros(1000, dtype=torch.float64), torch.zeros(1000, dtype=torch.float64)
    for k, gr in enumerate(grads, start=1):
        gd = gr.to(DEV).contiguous()
        assert L.psinn_adam(p(wd), p(gd), p(m), p(v), 1000, 1e-2, 0.9, 0.999, 1e-8, k, 1.0, None) == 0
        mr = 0.9 * mr + 0.1 * gr.double()
        vr = 0.999 * vr + 0.001 * gr.double() ** 2
        ref -= 1e-2 * (mr / (1 - 0.9 ** k)) / (torch.sqrt(vr / (1 - 0.999 ** k)) + 1e-8)
    torch.cuda.synchronize()
    np.testing.assert_allclose(wd.cpu().double().numpy(), ref.numpy(), rtol=1e-5, atol=1e-6)


def _small_solver(seed=4, n=40):
    b = basins.beta_plane_basin(seed=seed, nx=n, ny=n, n_islands=1)
    s = StreamfunctionSolver.from_basin(b, device=DEV)
    return b, s


def test_anorm_loss_and_gradient():
    b, s = _small_solver()
    inp = learned.SolverInputs(s)
    H, W = inp.H, inp.W
    rng = np.random.default_rng(0)
    Z = s.make_inv_sfc().cpu().numpy()
    act = (Z != 0)
    isl = islands_trace(b.kbot)
    cf = O.make_coeff(b)
    Wa = b.area_z()
    B = 2
    R, E, ER = [], [], []
    for _ in range(B):
        e = rng.normal(size=b.shape) * act
        r = O.apply_op(b.nx, b.ny, cf, e) * act
        R.append(r)
        E.append(e)
        ER.append((e * Wa * r)[act].sum())
    z = np.zeros((B, H, W), np.float32)
    z[:, :b.ny + 4, :b.nx + 4] = rng.normal(size=(B,) + b.shape)
    zd = torch.as_tensor(z, device=DEV)
    Rd = torch.as_tensor(np.array(R), device=DEV)
    ERd = torch.as_tensor(np.array(ER), device=DEV)
    work = torch.empty(2 * B, dtype=torch.float64, device=DEV)
    loss = torch.empty(1, dtype=torch.float64, device=DEV)
    gz = torch.empty_like(zd)
    assert s.L.psi_nn_anorm_loss(s.ctx, p(zd), H, W, B, p(Rd), p(ERd), p(work), p(gz), p(loss)) == 0
    torch.cuda.synchronize()
    # numpy FP64 reference
    Lref, Gref = 0.0, np.zeros((B, H, W))
    for k in range(B):
        zm = z[k, :b.ny + 4, :b.nx + 4].astype(np.float64) * act
        Sz = Wa * O.apply_op(b.nx, b.ny, cf, zm) * act
        a = (zm * Wa * R[k])[act].sum()
        q = (zm * Sz)[act].sum()
        Lref += (1 - a * a / (q * ER[k])) / B
        gk = (-2 * a / (q * ER[k]) * Wa * R[k] + 2 * a * a / (q * q * ER[k]) * Sz) * act / B
        Gref[k, :b.ny + 4, :b.nx + 4] = gk
    assert abs(float(loss) - Lref) <= 1e-4 * max(abs(Lref), 1e-3)
    assert cos(gz.cpu(), torch.as_tensor(Gref)) >= 0.9999


def test_training_reduces_loss_and_fcg_converges():
    b, s = _small_solver(seed=6, n=64)
    forc = torch.as_tensor(b.forc, device=DEV)
    rng = np.random.default_rng(1)
    ts = learned.make_trainset(s, forc, rng, n_synthetic=12, cg_iters=(5, 20, 60, 150))
    net = PrecondNet(device=DEV, seed=0)
    idx = list(range(min(8, ts.R.shape[0])))
    l0, _ = learned.loss_and_grad(net, ts, idx)
    l0 = float(l0)
    learned.train(net, [ts], steps=150, batch=4, lr=2e-3, log=None)
    l1, _ = learned.loss_and_grad(net, ts, idx)
    l1 = float(l1)
    print(f"  loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.8 * l0
    ls = learned.LearnedSolver(s, net)
    sol = s.zeros()
    its, conv, info = ls.solve(forc, sol, relres_target=1e-10, max_iterations=3000,
                               policy="learned")
    print(f"  learned FCG: {its} iterations, info={info}")
    assert conv and info["fallback_iterations"] == 0
    cf = O.make_coeff(b)
    isl = islands_trace(b.kbot)
    Z = O.make_inv(b.nx, b.ny, cf, O.Islands(isl.nisle, isl.offsets, isl.bi, isl.bj, isl.di,
                                             isl.dj, None))
    assert O.true_relres(b, cf, Z, b.forc, sol.cpu().numpy()) <= 1.01e-10
    # what the trained network gains over Jacobi-PCG on the same solve
    xj = s.zeros()
    jits, jconv, _ = s.congrad_streamfunction(forc, xj, mode="mixed", criterion="relres",
                                              relres_target=1e-10, max_iterations=20000)
    print(f"  Jacobi-PCG (mixed): {jits} iterations; learned FCG: {its}")
    assert jconv and its < jits
    # the measured policy probes both from one first guess, picks the one
    # with fewer ms per residual decade and converges
    sol2 = s.zeros()
    its2, conv2, info2 = ls.solve(forc, sol2, relres_target=1e-10, policy="auto")
    print(f"  auto policy: {info2['choice']} {its2} iterations, probe={info2['probe']}")
    assert conv2 and info2["choice"] in ("learned", "jacobi")
    pr = info2["probe"]
    assert info2["choice"] == ("learned" if pr["learned_ms_per_decade"] <
                               pr[f"{pr['fallback_mode']}_ms_per_decade"] else "jacobi")
    assert O.true_relres(b, cf, Z, b.forc, sol2.cpu().numpy()) <= 1.01e-10
