"""PyTorch fp32 restatement of the learned-guess U-Net (libmovfem_gpu.so,
mvf_nn_*), the checker its BF16 matrix-core kernels are compared against:
same layers, parameter layout and loss; weights and inputs rounded to BF16 as
the device holds them, activations kept in fp32."""
import numpy as np
import torch
import torch.nn.functional as F

CH = [16, 32, 64, 96]


def layer_shapes():
    out = []
    for k in range(4):
        cin = 16 if k == 0 else CH[k - 1]
        out += [(3, cin, CH[k]), (3, CH[k], CH[k])]
    for k in range(3):
        out += [(3, CH[k + 1] + CH[k], CH[k]), (3, CH[k], CH[k])]
    out.append((1, 16, 16))
    return out


def bf16(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def split(p):
    res, off = [], 0
    for ks, cin, cout in layer_shapes():
        nw = ks ** 3 * cin * cout
        w = p[off:off + nw].reshape(ks ** 3, cout, cin)
        off += nw
        b = p[off:off + cout]
        off += cout
        res.append((ks, w, b))
    assert off == p.size
    return res


def forward_backward(params, x, t, m, leaky=0.1):
    """loss, flat gradient (device layout), prediction (B, X, Y, Z, 12)."""
    ws = []
    for ks, w, b in split(params):
        wt = torch.from_numpy(bf16(w).reshape(ks, ks, ks, w.shape[1], w.shape[2]).copy()).permute(3, 4, 0, 1, 2)
        ws.append((ks, wt.contiguous().requires_grad_(True), torch.from_numpy(b.copy()).requires_grad_(True)))
    B, X, Y, Z, cin = x.shape
    xp = np.zeros((B, X, Y, Z, 16), np.float32)
    xp[..., :cin] = bf16(x)
    h = torch.from_numpy(xp).permute(0, 4, 1, 2, 3).contiguous()

    def conv(u, i):
        return F.conv3d(u, ws[i][1], ws[i][2], padding=ws[i][0] // 2)

    def act(u):
        return F.leaky_relu(u, leaky)

    skips = []
    for k in range(4):
        h = act(conv(h, 2 * k))
        h = act(conv(h, 2 * k + 1))
        if k < 3:
            skips.append(h)
            h = F.avg_pool3d(h, 2)
    for k in (2, 1, 0):
        h = torch.cat([F.interpolate(h, scale_factor=2, mode="nearest"), skips[k]], 1)
        h = act(conv(h, 8 + 2 * k))
        h = act(conv(h, 9 + 2 * k))
    out = conv(h, 14)[:, :12].permute(0, 2, 3, 4, 1)
    tt, mm = torch.from_numpy(t), torch.from_numpy(m)
    loss = (mm * (out - tt) ** 2).sum() / mm.sum()
    loss.backward()
    g = []
    for ks, w, b in ws:
        g += [w.grad.permute(2, 3, 4, 0, 1).reshape(ks ** 3, w.shape[0], w.shape[1]).numpy().ravel(), b.grad.numpy()]
    return float(loss.detach()), np.concatenate(g), out.detach().numpy()
