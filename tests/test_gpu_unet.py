"""The U-Net kernels of the learned preconditioner against a plain PyTorch float64 reference of
the same operations on the CPU (bf16 rounding of the kernels' outputs applied to the reference
where the kernels round):

- 3x3 convolution forward (both channel tilings), with and without the input mask; its bf16
  output may differ from the rounded reference by one bf16 ulp where the fp32 sums straddle a
  rounding boundary: relative L2 < 4e-3;
- the backward data pass (flipped weights) against autograd's input gradient, the weight and
  bias gradients against autograd's (fp32 accumulation of exact bf16 products: 1e-5);
- pooling, upsampling, channel copies, projection, loss (exact or to fp32 rounding);
- the whole network: output and every parameter gradient against autograd through the
  reference network (2e-2 / 5e-2: the kernels carry activation gradients in bf16), and a few
  Adam steps lower the loss.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_CONV = 4e-3


def _t():
    import torch
    return torch


def bf(x):
    torch = _t()
    return x.to(torch.bfloat16).to(torch.float64)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev(x, dtype):
    torch = _t()
    return torch.as_tensor(x).to(dtype).to('cuda').contiguous()


def _lib():
    from stride_amd import _native
    return _native.lib()


def _st():
    return None  # the null stream: every launch is followed by a synchronising copy here


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _ref_conv(x, w, b, relu):
    """x (N,H,W,Cin), w (3,3,Cout,Cin) -> (N,H,W,Cout) in float64."""
    torch = _t()
    import torch.nn.functional as F
    y = F.conv2d(x.permute(0, 3, 1, 2), w.permute(2, 3, 0, 1), bias=b, padding=1).permute(0, 2, 3, 1)
    return torch.relu(y) if relu else y


@pytest.mark.parametrize('cin,cout,masked', [(64, 96, False), (32, 64, True), (96, 32, True)])
def test_conv3x3_forward(gpu, cin, cout, masked):
    torch = _t()
    N, H, W = 2, 8, 128
    rng = np.random.default_rng(cin + cout)
    x = bf(torch.tensor(rng.standard_normal((N, H, W, cin))))
    w = bf(torch.tensor(rng.standard_normal((3, 3, cout, cin)) * 0.1))
    b = torch.tensor(rng.standard_normal(cout) * 0.1).float().double()
    m = bf(torch.tensor(rng.standard_normal((N, H, W, cin)))) if masked else None
    print('conv3x3 %d -> %d masked=%s' % (cin, cout, masked), flush=True)
    y = torch.empty((N, H, W, cout), dtype=torch.bfloat16, device='cuda')
    # device copies held in locals: a temporary freed before the launch would hand its memory
    # to the next upload
    xd, wd, bd = _dev(x, torch.bfloat16), _dev(w, torch.bfloat16), _dev(b, torch.float32)
    md = _dev(m, torch.bfloat16) if masked else None
    rc = _lib().sa_nn_conv3x3(_st(), _p(xd), _p(md), _p(wd), _p(bd), _p(y), N, H, W, cin, cout, 1)
    assert rc == 0
    xin = x * (m > 0) if masked else x
    ref = bf(_ref_conv(xin, w, b, True))
    assert rel(y.double().cpu(), ref) < TOL_CONV


def test_conv_backward_and_weight_gradients(gpu):
    torch = _t()
    N, H, W, cin, cout = 2, 8, 64, 64, 128
    rng = np.random.default_rng(5)
    x = bf(torch.tensor(rng.standard_normal((N, H, W, cin))))
    w = bf(torch.tensor(rng.standard_normal((3, 3, cout, cin)) * 0.1))
    y = bf(torch.tensor(rng.standard_normal((N, H, W, cout))))         # the layer output (mask)
    dy = bf(torch.tensor(rng.standard_normal((N, H, W, cout))))
    dpre = dy * (y > 0)
    xr = x.clone().requires_grad_(True)
    wr = w.clone().requires_grad_(True)
    br = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    out = _ref_conv(xr, wr, br, False)
    (out * dpre).sum().backward()
    print('conv backward %d -> %d' % (cin, cout), flush=True)
    L = _lib()
    dx = torch.empty((N, H, W, cin), dtype=torch.bfloat16, device='cuda')
    wf = torch.empty((3, 3, cin, cout), dtype=torch.bfloat16, device='cuda')
    wd, yd, dyd, xd = (_dev(v, torch.bfloat16) for v in (w, y, dy, x))
    assert L.sa_nn_flip_weights(_st(), _p(wd), _p(wf), cin, cout) == 0
    assert L.sa_nn_conv3x3(_st(), _p(dyd), _p(yd), _p(wf), None, _p(dx), N, H, W, cout, cin, 0) == 0
    assert rel(dx.double().cpu(), bf(xr.grad)) < TOL_CONV
    npx = N * H * W
    for pps in (64, 2048):
        nsplit = (npx + pps - 1) // pps
        part = torch.empty(nsplit * 9 * cin * cout, dtype=torch.float32, device='cuda')
        dw = torch.full((3, 3, cout, cin), 5.0, dtype=torch.float32, device='cuda')
        assert L.sa_nn_wgrad3x3(_st(), _p(dyd), _p(yd), _p(xd), _p(part), _p(dw), N, H, W, cin, cout, pps, 0) == 0
        assert rel(dw.double().cpu(), wr.grad) < 1e-5
        # accumulate adds to what is there
        assert L.sa_nn_wgrad3x3(_st(), _p(dyd), _p(yd), _p(xd), _p(part), _p(dw), N, H, W, cin, cout, pps, 1) == 0
        assert rel(dw.double().cpu(), 2 * wr.grad) < 1e-5
    bpart = torch.empty(((npx + 99) // 100) * cout, dtype=torch.float32, device='cuda')
    db = torch.empty(cout, dtype=torch.float32, device='cuda')
    assert L.sa_nn_bias_grad(_st(), _p(dyd), _p(yd), _p(bpart), _p(db), ctypes.c_longlong(npx), cout, 100, 0) == 0
    assert rel(db.double().cpu(), br.grad) < 1e-6


def test_pool_upsample_copy_head_loss(gpu):
    torch = _t()
    import torch.nn.functional as F
    L = _lib()
    N, H, W, C = 2, 8, 64, 32
    rng = np.random.default_rng(9)
    x = bf(torch.tensor(rng.standard_normal((N, H, W, C))))
    xd = _dev(x, torch.bfloat16)
    y = torch.empty((N, H // 2, W // 2, C), dtype=torch.bfloat16, device='cuda')
    assert L.sa_nn_avgpool(_st(), _p(xd), _p(y), N, H, W, C) == 0
    ref = bf(F.avg_pool2d(x.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1))
    assert rel(y.double().cpu(), ref) < 1e-6
    dx = _dev(x, torch.bfloat16)
    assert L.sa_nn_avgpool_back(_st(), _p(y), _p(dx), N, H, W, C, 1) == 0
    want = bf(x + 0.25 * ref.repeat_interleave(2, 1).repeat_interleave(2, 2))
    assert rel(dx.double().cpu(), want) < 1e-6
    cat = torch.zeros((N, 2 * H, 2 * W, 3 * C), dtype=torch.bfloat16, device='cuda')
    big = _dev(bf(torch.tensor(rng.standard_normal((N, 2 * H, 2 * W, C)))), torch.bfloat16)
    assert L.sa_nn_copy_channels(_st(), _p(big), C, 0, _p(cat), 3 * C, 0, C, ctypes.c_longlong(N * 4 * H * W), 0) == 0
    xx = torch.cat([xd, xd], -1).contiguous()
    assert L.sa_nn_upsample(_st(), _p(xx), _p(cat), N, H, W, 2 * C, 3 * C, C) == 0
    up = xx.double().cpu().repeat_interleave(2, 1).repeat_interleave(2, 2)
    assert torch.equal(cat[..., :C].cpu(), big.cpu()) and torch.equal(cat[..., C:].double().cpu(), up)
    back = torch.empty((N, H, W, 2 * C), dtype=torch.bfloat16, device='cuda')
    assert L.sa_nn_upsample_back(_st(), _p(cat), _p(back), N, H, W, 2 * C, 3 * C, C) == 0
    assert rel(back.double().cpu(), bf(4 * xx.double().cpu())) < 1e-6
    # projection and loss
    hw = torch.tensor(rng.standard_normal(C), dtype=torch.float32)
    hb = torch.tensor([0.3], dtype=torch.float32)
    hwd, hbd = hw.cuda(), hb.cuda()
    out = torch.empty((N, H, W), dtype=torch.float32, device='cuda')
    assert L.sa_nn_head(_st(), _p(xd), _p(hwd), _p(hbd), _p(out), ctypes.c_longlong(N * H * W), C) == 0
    ref = (x * hw.double()).sum(-1) + 0.3
    assert rel(out.double().cpu(), ref) < 1e-6
    t = torch.tensor(rng.standard_normal((N, H, W)), dtype=torch.float32)
    n = N * H * W
    part = torch.empty(n // 256 + 2, dtype=torch.float64, device='cuda')
    d = torch.empty_like(out)
    loss = ctypes.c_double()
    td = t.cuda()
    assert L.sa_nn_mse(_st(), _p(out), _p(td), _p(d), ctypes.c_longlong(n), ctypes.c_float(1 / n), _p(part),
                       ctypes.byref(loss)) == 0
    r = out.double().cpu() - t.double()
    assert loss.value == pytest.approx(0.5 * float((r ** 2).mean()), rel=1e-6)
    assert rel(d.double().cpu(), r / n) < 1e-6


def _ref_network(net, g, scale, model=None, model_scale=1.0):
    """The same network in float64 on the CPU from the network's parameters (bf16-rounded
    weights, activations rounded to bf16 where the kernels round)."""
    torch = _t()
    import torch.nn.functional as F
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in net.state().items()}

    def conv(x, name):
        return bf(_ref_conv(x, bf(P[name + '.w']), P[name + '.b'].float().double(), True))

    def pool(x):
        return bf(F.avg_pool2d(x.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1))

    def up(x):
        return x.repeat_interleave(2, 1).repeat_interleave(2, 2)

    N, H, W = g.shape
    x0 = torch.zeros((N, H, W, 32), dtype=torch.float64)
    x0[..., 0] = bf(scale * g)
    if model is not None:
        x0[..., 1] = bf(model_scale * model)
    e1 = conv(conv(x0, 'e1a'), 'e1')
    e2 = conv(conv(pool(e1), 'e2a'), 'e2')
    e3 = conv(conv(pool(e2), 'e3a'), 'e3')
    b = conv(conv(pool(e3), 'ba'), 'b')
    d3 = conv(conv(torch.cat([e3, up(b)], -1), 'd3a'), 'd3')
    d2 = conv(conv(torch.cat([e2, up(d3)], -1), 'd2a'), 'd2')
    d1 = conv(conv(torch.cat([e1, up(d2)], -1), 'd1a'), 'd1')
    out = (d1 * P['head.w']).sum(-1) + P['head.b']
    return out, P


def test_unet_forward_backward_against_reference(gpu):
    torch = _t()
    from stride_amd.unet import UNet
    net = UNet(device=0, seed=3)
    rng = np.random.default_rng(11)
    N, H, W = 1, 32, 512
    g = torch.tensor(rng.standard_normal((N, H, W)), dtype=torch.float32)
    target = torch.tensor(rng.standard_normal((N, H, W)), dtype=torch.float32) * 0.1
    print('unet forward/backward %s' % ((N, H, W),), flush=True)
    model = torch.tensor(rng.standard_normal((N, H, W)), dtype=torch.float32)
    out = net.forward(g.cuda(), scale=0.5, model=model.cuda(), model_scale=0.25)
    loss, d = net.mse(out, target.cuda())
    net.backward(d)
    ref_out, P = _ref_network(net, g.double(), 0.5, model.double(), 0.25)
    ref_loss = 0.5 * ((ref_out - target.double()) ** 2).mean()
    ref_loss.backward()
    assert rel(out.cpu(), ref_out.detach()) < 2e-2
    assert loss == pytest.approx(float(ref_loss), rel=2e-2)
    for name, prm in net.params.items():
        if name == 'e1a.w':
            got, want = prm.g.cpu().numpy()[..., :2], P[name].grad.numpy()[..., :2]  # the live input channels
        else:
            got, want = prm.g.cpu().numpy(), P[name].grad.numpy()
        assert rel(got, want) < 5e-2, name


def test_unet_training_lowers_loss(gpu):
    torch = _t()
    from stride_amd.unet import UNet
    net = UNet(device=0, seed=1)
    rng = np.random.default_rng(2)
    N, H, W = 2, 32, 512
    g = torch.tensor(rng.standard_normal((N, H, W)), dtype=torch.float32, device='cuda')
    target = 0.5 * g.clone()     # a map the network can learn
    losses = [net.train_step(g, target, scale=1.0, lr=3e-3) for _ in range(30)]
    print('unet training losses %.4g -> %.4g' % (losses[0], losses[-1]), flush=True)
    assert losses[-1] < 0.3 * losses[0]


def test_saved_network_reloads(gpu, tmp_path):
    """save_network / load_network: the reloaded network (and its input scales) gives the same
    output bit for bit."""
    torch = _t()
    from stride_amd.unet import UNet
    from stride_amd import learned
    net = UNet(device=0, seed=3)
    rng = np.random.default_rng(4)
    g = torch.tensor(rng.standard_normal((2, 32, 512)), dtype=torch.float32, device='cuda')
    net.train_step(g, 0.3 * g, scale=1.0, lr=1e-3)
    path = str(tmp_path / 'net.npz')
    learned.save_network(net, {300e3: 0.5, 700e3: 2.0}, path)
    net2, scales = learned.load_network(path)
    assert scales == {300e3: 0.5, 700e3: 2.0}
    a, b = net.forward(g), net2.forward(g)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
