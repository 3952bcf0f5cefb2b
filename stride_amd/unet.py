"""The learned gradient preconditioner's network: a U-Net run by the library's own kernels.

Four resolution levels (R, R/2, R/4, R/8) with 32, 64, 128 and 256 channels; two 3x3
convolutions with ReLU per level on the way down, 2 x 2 average pooling between levels, nearest
x2 upsampling on the way up, the encoder's output of the same level concatenated in front of the
upsampled channels, two 3x3 convolutions with ReLU per level, and a 1 x 1 projection to one
output channel. Weights and activations are bf16 and every product accumulates in fp32 on the
matrix cores (``sa_nn_conv3x3`` / ``sa_nn_wgrad3x3``, v_mfma_f32_16x16x32_bf16); fp32 master
weights, Adam moments and the loss stay fp32. The input has two channels (the gradient and,
optionally, the current model), padded to 32 with zero channels so the first convolution runs
on the same kernel.

The network has no counterpart in the reference. It maps the raw vp gradient of an FWI
iteration (``get_grad_vp``, stride/physics/iso_acoustic/devito.py:1045-1111) to a model update;
``stride_amd.learned`` trains it on solver-generated pairs and keeps its update only when the
forward solver's misfit drops.

torch supplies the device memory (tensors) and the stream; every computation runs in the HIP
library, which must be built (``_native.lib()`` raises otherwise).
"""

import ctypes

import numpy as np

from . import _native

__all__ = ['UNet', 'LEVEL_CHANNELS']

LEVEL_CHANNELS = (32, 64, 128, 256)
_PX_PER_SPLIT = 2048   # pixels per workgroup share of the weight-gradient GEMM
_PX_PER_BLOCK = 1024   # pixels per partial sum of the bias and projection gradients


def _torch():
    import torch
    return torch


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check(rc):
    if rc != 0:
        raise _native.SolverError(rc, _native.lib().sa_last_error().decode())


class _Param:
    """fp32 master, bf16 copy (for 3x3 weights), gradient and Adam moments of one tensor."""

    def __init__(self, value, device, bf16):
        torch = _torch()
        self.w = torch.tensor(value, dtype=torch.float32, device=device)
        self.wb = self.w.to(torch.bfloat16) if bf16 else None
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)


class UNet:
    """U-Net on (N, H, W) fp32 images, H a multiple of 32 and W of 512 (every level's width a
    multiple of 64). ``forward`` keeps the activations for ``backward``."""

    # (name, input channels, output channels, level)
    def __init__(self, device=0, seed=0):
        torch = _torch()
        self.device = torch.device('cuda', device) if isinstance(device, int) else device
        self.lib = _native.lib()
        rng = np.random.default_rng(seed)
        c1, c2, c3, c4 = LEVEL_CHANNELS
        self.convs = [('e1a', c1, c1), ('e1', c1, c1), ('e2a', c1, c2), ('e2', c2, c2), ('e3a', c2, c3),
                      ('e3', c3, c3), ('ba', c3, c4), ('b', c4, c4), ('d3a', c3 + c4, c3), ('d3', c3, c3),
                      ('d2a', c2 + c3, c2), ('d2', c2, c2), ('d1a', c1 + c2, c1), ('d1', c1, c1)]
        self.params = {}
        for name, cin, cout in self.convs:
            fan_in = 9 * (2 if name == 'e1a' else cin)          # the padded input has two live channels
            w = rng.standard_normal((3, 3, cout, cin)) * np.sqrt(2.0 / fan_in)
            if name == 'e1a':
                w[:, :, :, 2:] = 0
            self.params[name + '.w'] = _Param(w, self.device, True)
            self.params[name + '.b'] = _Param(np.zeros(cout), self.device, False)
        self.params['head.w'] = _Param(rng.standard_normal(c1) * np.sqrt(1.0 / c1) * 0.1, self.device, False)
        self.params['head.b'] = _Param(np.zeros(1), self.device, False)
        self.step_count = 0
        self._cache = None

    # ---- helpers ------------------------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    def _empty(self, *shape, dtype=None):
        torch = _torch()
        return torch.empty(shape, dtype=dtype or torch.bfloat16, device=self.device)

    def _conv(self, x, name, N, H, W, relu=True):
        pw, pb = self.params[name + '.w'], self.params[name + '.b']
        cout, cin = pw.w.shape[2], pw.w.shape[3]
        y = self._empty(N, H, W, cout)
        _check(self.lib.sa_nn_conv3x3(self._stream(), _p(x), None, _p(pw.wb), _p(pb.w), _p(y), N, H, W, cin, cout,
                                      int(relu)))
        return y

    def _pool(self, x, N, H, W, C):
        y = self._empty(N, H // 2, W // 2, C)
        _check(self.lib.sa_nn_avgpool(self._stream(), _p(x), _p(y), N, H, W, C))
        return y

    def _cat_up(self, skip, low, N, H, W, cs, cl):
        """[skip | upsample(low)] along channels at H x W."""
        y = self._empty(N, H, W, cs + cl)
        npx = ctypes.c_longlong(N * H * W)
        _check(self.lib.sa_nn_copy_channels(self._stream(), _p(skip), cs, 0, _p(y), cs + cl, 0, cs, npx, 0))
        _check(self.lib.sa_nn_upsample(self._stream(), _p(low), _p(y), N, H // 2, W // 2, cl, cs + cl, cs))
        return y

    # ---- forward ------------------------------------------------------------------------------
    def forward(self, g, scale=1.0, model=None, model_scale=1.0):
        """g (and the optional second input channel `model`): (N, H, W) fp32 device tensors ->
        (N, H, W) fp32 output."""
        torch = _torch()
        N, H, W = g.shape
        if H % 32 or W % 512:
            raise ValueError('U-Net input must be (N, H, W) with H % 32 == 0 and W % 512 == 0')
        c1, c2, c3, c4 = LEVEL_CHANNELS
        g = g.contiguous().to(torch.float32)
        m = None if model is None else model.contiguous().to(torch.float32)
        x0 = self._empty(N, H, W, c1)
        _check(self.lib.sa_nn_pack_input(self._stream(), _p(g), ctypes.c_float(scale), _p(m),
                                         ctypes.c_float(model_scale), _p(x0), ctypes.c_longlong(N * H * W), c1))
        a = dict(x0=x0)
        a['e1a'] = self._conv(x0, 'e1a', N, H, W)
        a['e1'] = self._conv(a['e1a'], 'e1', N, H, W)
        a['p1'] = self._pool(a['e1'], N, H, W, c1)
        H2, W2 = H // 2, W // 2
        a['e2a'] = self._conv(a['p1'], 'e2a', N, H2, W2)
        a['e2'] = self._conv(a['e2a'], 'e2', N, H2, W2)
        a['p2'] = self._pool(a['e2'], N, H2, W2, c2)
        H3, W3 = H2 // 2, W2 // 2
        a['e3a'] = self._conv(a['p2'], 'e3a', N, H3, W3)
        a['e3'] = self._conv(a['e3a'], 'e3', N, H3, W3)
        a['p3'] = self._pool(a['e3'], N, H3, W3, c3)
        H4, W4 = H3 // 2, W3 // 2
        a['ba'] = self._conv(a['p3'], 'ba', N, H4, W4)
        a['b'] = self._conv(a['ba'], 'b', N, H4, W4)
        a['c3'] = self._cat_up(a['e3'], a['b'], N, H3, W3, c3, c4)
        a['d3a'] = self._conv(a['c3'], 'd3a', N, H3, W3)
        a['d3'] = self._conv(a['d3a'], 'd3', N, H3, W3)
        a['c2'] = self._cat_up(a['e2'], a['d3'], N, H2, W2, c2, c3)
        a['d2a'] = self._conv(a['c2'], 'd2a', N, H2, W2)
        a['d2'] = self._conv(a['d2a'], 'd2', N, H2, W2)
        a['c1'] = self._cat_up(a['e1'], a['d2'], N, H, W, c1, c2)
        a['d1a'] = self._conv(a['c1'], 'd1a', N, H, W)
        a['d1'] = self._conv(a['d1a'], 'd1', N, H, W)
        out = torch.empty((N, H, W), dtype=torch.float32, device=self.device)
        _check(self.lib.sa_nn_head(self._stream(), _p(a['d1']), _p(self.params['head.w'].w),
                                   _p(self.params['head.b'].w), _p(out), ctypes.c_longlong(N * H * W), c1))
        self._cache = (a, (N, H, W))
        return out

    __call__ = forward

    # ---- backward -----------------------------------------------------------------------------
    def _conv_back(self, dy, name, x, y, N, H, W, need_dx=True):
        """Gradients of y = relu(conv(x)) given dL/dy: weight and bias gradients into the
        parameter's .g, and dL/dx (None if not needed)."""
        torch = _torch()
        pw, pb = self.params[name + '.w'], self.params[name + '.b']
        cout, cin = pw.w.shape[2], pw.w.shape[3]
        st = self._stream()
        npx = N * H * W
        nsplit = (npx + _PX_PER_SPLIT - 1) // _PX_PER_SPLIT
        part = torch.empty(nsplit * 9 * cin * cout, dtype=torch.float32, device=self.device)
        _check(self.lib.sa_nn_wgrad3x3(st, _p(dy), _p(y), _p(x), _p(part), _p(pw.g), N, H, W, cin, cout,
                                       _PX_PER_SPLIT, 0))
        nb = (npx + _PX_PER_BLOCK - 1) // _PX_PER_BLOCK
        bpart = torch.empty(nb * cout, dtype=torch.float32, device=self.device)
        _check(self.lib.sa_nn_bias_grad(st, _p(dy), _p(y), _p(bpart), _p(pb.g), ctypes.c_longlong(npx), cout,
                                        _PX_PER_BLOCK, 0))
        if not need_dx:
            return None
        wf = self._empty(3, 3, cin, cout)
        _check(self.lib.sa_nn_flip_weights(st, _p(pw.wb), _p(wf), cin, cout))
        dx = self._empty(N, H, W, cin)
        _check(self.lib.sa_nn_conv3x3(st, _p(dy), _p(y), _p(wf), None, _p(dx), N, H, W, cout, cin, 0))
        return dx

    def _split_up(self, dc, d_skip_into, N, H, W, cs, cl):
        """Backward of [skip | upsample(low)]: the skip part into d_skip_into (overwritten), the
        low part summed over its children (returned, at H/2 x W/2)."""
        st = self._stream()
        _check(self.lib.sa_nn_copy_channels(st, _p(dc), cs + cl, 0, _p(d_skip_into), cs, 0, cs,
                                            ctypes.c_longlong(N * H * W), 0))
        d_low = self._empty(N, H // 2, W // 2, cl)
        _check(self.lib.sa_nn_upsample_back(st, _p(dc), _p(d_low), N, H // 2, W // 2, cl, cs + cl, cs))
        return d_low

    def backward(self, dout):
        """dout: (N, H, W) fp32 dL/d(output) of the last forward. Fills every parameter's .g."""
        torch = _torch()
        a, (N, H, W) = self._cache
        c1, c2, c3, c4 = LEVEL_CHANNELS
        st = self._stream()
        H2, W2, H3, W3, H4, W4 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        npx = N * H * W
        nb = (npx + _PX_PER_BLOCK - 1) // _PX_PER_BLOCK
        hpart = torch.empty(nb * (c1 + 1), dtype=torch.float32, device=self.device)
        hg = torch.empty(c1 + 1, dtype=torch.float32, device=self.device)
        d_d1 = self._empty(N, H, W, c1)
        _check(self.lib.sa_nn_head_back(st, _p(dout.contiguous()), _p(self.params['head.w'].w), _p(a['d1']),
                                        _p(d_d1), _p(hpart), _p(hg), ctypes.c_longlong(npx), c1, _PX_PER_BLOCK, 0))
        self.params['head.w'].g.copy_(hg[:c1])
        self.params['head.b'].g.copy_(hg[c1:])
        d = self._conv_back(d_d1, 'd1', a['d1a'], a['d1'], N, H, W)
        d_c1 = self._conv_back(d, 'd1a', a['c1'], a['d1a'], N, H, W)
        d_e1 = self._empty(N, H, W, c1)
        d_d2 = self._split_up(d_c1, d_e1, N, H, W, c1, c2)
        d = self._conv_back(d_d2, 'd2', a['d2a'], a['d2'], N, H2, W2)
        d_c2 = self._conv_back(d, 'd2a', a['c2'], a['d2a'], N, H2, W2)
        d_e2 = self._empty(N, H2, W2, c2)
        d_d3 = self._split_up(d_c2, d_e2, N, H2, W2, c2, c3)
        d = self._conv_back(d_d3, 'd3', a['d3a'], a['d3'], N, H3, W3)
        d_c3 = self._conv_back(d, 'd3a', a['c3'], a['d3a'], N, H3, W3)
        d_e3 = self._empty(N, H3, W3, c3)
        d_b = self._split_up(d_c3, d_e3, N, H3, W3, c3, c4)
        d = self._conv_back(d_b, 'b', a['ba'], a['b'], N, H4, W4)
        d_p3 = self._conv_back(d, 'ba', a['p3'], a['ba'], N, H4, W4)
        _check(self.lib.sa_nn_avgpool_back(st, _p(d_p3), _p(d_e3), N, H3, W3, c3, 1))
        d = self._conv_back(d_e3, 'e3', a['e3a'], a['e3'], N, H3, W3)
        d_p2 = self._conv_back(d, 'e3a', a['p2'], a['e3a'], N, H3, W3)
        _check(self.lib.sa_nn_avgpool_back(st, _p(d_p2), _p(d_e2), N, H2, W2, c2, 1))
        d = self._conv_back(d_e2, 'e2', a['e2a'], a['e2'], N, H2, W2)
        d_p1 = self._conv_back(d, 'e2a', a['p1'], a['e2a'], N, H2, W2)
        _check(self.lib.sa_nn_avgpool_back(st, _p(d_p1), _p(d_e1), N, H, W, c1, 1))
        d = self._conv_back(d_e1, 'e1', a['e1a'], a['e1'], N, H, W)
        self._conv_back(d, 'e1a', a['x0'], a['e1a'], N, H, W, need_dx=False)

    # ---- loss and optimiser -------------------------------------------------------------------
    def mse(self, out, target, want_grad=True):
        """(0.5 mean((out - target)^2), d loss / d out or None)."""
        torch = _torch()
        n = out.numel()
        part = torch.empty((n + 255) // 256 + 1, dtype=torch.float64, device=self.device)
        d = torch.empty_like(out) if want_grad else None
        loss = ctypes.c_double()
        _check(self.lib.sa_nn_mse(self._stream(), _p(out), _p(target.contiguous()), _p(d), ctypes.c_longlong(n),
                                  ctypes.c_float(1.0 / n), _p(part), ctypes.byref(loss)))
        return loss.value, d

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        st = self._stream()
        for prm in self.params.values():
            _check(self.lib.sa_nn_adam(st, _p(prm.w), _p(prm.g), _p(prm.m), _p(prm.v), _p(prm.wb),
                                       ctypes.c_longlong(prm.w.numel()), ctypes.c_float(lr), ctypes.c_float(b1),
                                       ctypes.c_float(b2), ctypes.c_float(eps), self.step_count))
        # the padded input channels of the first layer stay zero
        # (their gradient is zero: the padded input channels are zero)

    def train_step(self, g, target, scale=1.0, lr=1e-3, model=None, model_scale=1.0):
        out = self.forward(g, scale, model, model_scale)
        loss, d = self.mse(out, target)
        self.backward(d)
        self.adam_step(lr)
        return loss

    def state(self):
        """Host copies of the fp32 master parameters."""
        return {k: v.w.detach().cpu().numpy() for k, v in self.params.items()}

    def load_state(self, state):
        torch = _torch()
        for k, v in state.items():
            prm = self.params[k]
            prm.w.copy_(torch.as_tensor(v, dtype=torch.float32))
            if prm.wb is not None:
                prm.wb.copy_(prm.w.to(torch.bfloat16))
