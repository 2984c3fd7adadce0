"""CPU reference of the denoiser network for the tests: the same U-Net in
torch float32 on the CPU (torch.nn.functional.conv3d and autograd), with every
stored activation rounded to bfloat16 as the GPU stores them.  The weights
are rounded to bfloat16 for the convolutions (the GPU packs BF16 copies of
its FP32 master weights); biases and the head stay FP32."""
import numpy as np
import torch
import torch.nn.functional as F

from tam_amd.denoise import LAYERS, param_layout


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def unpack(p):
    """{name: float32 tensor} from the flat parameter vector."""
    out = {}
    for name, off, shape in param_layout():
        out[name] = torch.tensor(np.asarray(p[off:off + int(np.prod(shape))], dtype=np.float32).reshape(shape))
    return out


def _conv(x, w, b):
    # w: [27][cin][cout], tap = 9 (dz+1) + 3 (dy+1) + (dx+1) -> (cout, cin, 3, 3, 3)
    cin, cout = w.shape[1], w.shape[2]
    wt = bf(w).reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2).contiguous()
    return bf(F.relu(F.conv3d(x, wt, bias=b, padding=1)))


def forward(params, features):
    """features (4, nz, ny, nx) -> prediction (nz, ny, nx); params: unpack()."""
    f = torch.as_tensor(np.asarray(features, dtype=np.float32))
    x0 = torch.zeros((1, 16) + tuple(f.shape[1:]))
    x0[0, :4] = bf(f)
    a0 = _conv(x0, params["w0"], params["b0"])
    a1 = _conv(a0, params["w1"], params["b1"])
    p1 = bf(F.avg_pool3d(a1, 2))
    a2 = _conv(p1, params["w2"], params["b2"])
    a3 = _conv(a2, params["w3"], params["b3"])
    p3 = bf(F.avg_pool3d(a3, 2))
    a4 = _conv(p3, params["w4"], params["b4"])
    a5 = _conv(a4, params["w5"], params["b5"])
    c6 = torch.cat([F.interpolate(a5, scale_factor=2, mode="nearest"), a3], dim=1)
    a6 = _conv(c6, params["w6"], params["b6"])
    a7 = _conv(a6, params["w7"], params["b7"])
    c8 = torch.cat([F.interpolate(a7, scale_factor=2, mode="nearest"), a1], dim=1)
    a8 = _conv(c8, params["w8"], params["b8"])
    a9 = _conv(a8, params["w9"], params["b9"])
    head = torch.einsum("c,ncdhw->ndhw", params["head_w"], a9) + params["head_b"][0]
    return (x0[:, 0] + head)[0]


def gradients(p, features, target):
    """(flat gradient of mean |pred - target|, loss) by autograd."""
    params = unpack(p)
    for t in params.values():
        t.requires_grad_(True)
    pred = forward(params, features)
    loss = (pred - torch.as_tensor(np.asarray(target, dtype=np.float32))).abs().mean()
    loss.backward()
    g = np.concatenate([params[name].grad.numpy().ravel() for name, _, _ in param_layout()])
    return g, float(loss.detach())


def random_params(n, seed, scale=1.0):
    """Parameters with every layer active (He scale weights, small biases,
    nonzero head)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    for name, off, shape in param_layout():
        size = int(np.prod(shape))
        if name.startswith("w"):
            i = int(name[1:])
            fan = 27 * (4 if i == 0 else LAYERS[i][0])
            v = rng.normal(0.0, scale * np.sqrt(2.0 / fan), size)
            if i == 0:
                v = v.reshape(shape)
                v[:, 4:, :] = 0.0
                v = v.ravel()
        elif name.startswith("b"):
            v = rng.normal(0.0, 0.05, size)
        elif name == "head_w":
            v = rng.normal(0.0, 0.2, size)
        else:
            v = np.array([0.01])
        p[off:off + size] = v
    return p
