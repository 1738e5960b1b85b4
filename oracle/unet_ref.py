"""CPU restatement of the learned-proposal U-Net of tom_amd/csrc/unet.hip.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP U-Net
(there is no reference counterpart for this network). It is never part of the
product path.

The network (layer order = parameter order of tomhip_unet_get_params):
  e1a 16->16, e1b 16->16 (skip1), 2x2x2 average pool,
  e2a 16->32, e2b 32->32 (skip2), average pool,
  b1 32->64, b2 64->64, up x2 (trilinear, half-pixel, clamped),
  d2a [up(b2) | skip2] 96->32, d2b 32->32, up x2,
  d1a [up(d2b) | skip1] 48->16, d1b 16->16,
  head 1x1x1 16->80 (channel 0 particle logit, 1..nbin orientation bins).
3x3x3 convolutions, zero padding, ReLU except the head. The input volume sits
in channel 0 of 16, an optional auxiliary map in channel 1 (the other
channels are zero).
Parameters per layer: W[t][co][ci] (t = 9(dz+1) + 3(dy+1) + (dx+1), output
voxel v reads input v + (dx, dy, dz)) followed by bias[co].

`bf16=True` rounds to bfloat16 (nearest-even) where the HIP path stores or
feeds bfloat16: the input, every weight, and every stored activation; the head
output stays float.

Parity status: the tap convention is pinned by a brute-force loop
(conv_direct) and the gradients by central finite differences in FP64.
"""
import numpy as np
import torch
import torch.nn.functional as F

SPEC = [(16, 16, 27, 1), (16, 16, 27, 1), (16, 32, 27, 1), (32, 32, 27, 1), (32, 64, 27, 1), (64, 64, 27, 1),
        (96, 32, 27, 1), (32, 32, 27, 1), (48, 16, 27, 1), (16, 16, 27, 1), (16, 80, 1, 0)]
HEAD = 80


def num_params():
    return sum(t * co * ci + co for ci, co, t, _ in SPEC)


def bf16_round(a):
    """Round float32/float64 values to bfloat16 (nearest-even), returned as the
    same dtype."""
    t = torch.as_tensor(a)
    return t.to(torch.float32).to(torch.bfloat16).to(t.dtype)


def unpack(theta, dtype=torch.float64):
    """flat parameter vector -> list of (weight [Cout][Cin][kz][ky][kx], bias)."""
    theta = torch.as_tensor(np.asarray(theta), dtype=dtype)
    out, off = [], 0
    for ci, co, t, _ in SPEC:
        nw = t * co * ci
        W = theta[off:off + nw].reshape(t, co, ci)
        b = theta[off + nw:off + nw + co]
        off += nw + co
        k = 3 if t == 27 else 1
        # W[t][co][ci] with t = 9 kz + 3 ky + kx  ->  torch [co][ci][kz][ky][kx]
        w = W.permute(1, 2, 0).reshape(co, ci, k, k, k)
        out.append((w, b))
    return out


def forward(vol, theta, bf16=True, params=None, aux=None):
    """vol (and the optional aux map): [D][H][W] arrays (z, y, x). Returns
    logits [D][H][W][80] and the parameter tensors (for autograd)."""
    P = params if params is not None else unpack(theta)
    q = bf16_round if bf16 else (lambda a: a)
    x = torch.as_tensor(np.asarray(vol), dtype=torch.float64)
    x0 = torch.zeros((1, 16) + tuple(x.shape), dtype=torch.float64)
    x0[0, 0] = q(x)
    if aux is not None:
        x0[0, 1] = q(torch.as_tensor(np.asarray(aux), dtype=torch.float64))

    def conv(i, inp):
        w, b = P[i]
        wq = w + (q(w.detach()) - w.detach())  # bf16 value, identity gradient
        pad = 1 if SPEC[i][2] == 27 else 0
        y = F.conv3d(inp, wq, b, padding=pad)
        if SPEC[i][3]:
            y = F.relu(y)
        return y

    def st(a):  # stored activation: bf16 value, identity gradient
        return a + (q(a.detach()) - a.detach())

    a1 = st(conv(0, x0))
    s1 = st(conv(1, a1))
    p1 = st(F.avg_pool3d(s1, 2))
    a2 = st(conv(2, p1))
    s2 = st(conv(3, a2))
    p2 = st(F.avg_pool3d(s2, 2))
    b1 = st(conv(4, p2))
    b2 = st(conv(5, b1))
    u2 = st(F.interpolate(b2, scale_factor=2, mode="trilinear", align_corners=False))
    d2a = st(conv(6, torch.cat([u2, s2], 1)))
    d2b = st(conv(7, d2a))
    u1 = st(F.interpolate(d2b, scale_factor=2, mode="trilinear", align_corners=False))
    d1a = st(conv(8, torch.cat([u1, s1], 1)))
    d1b = st(conv(9, d1a))
    logits = conv(10, d1b)
    return logits[0].permute(1, 2, 3, 0), P


def loss(logits, tprob, tbin, nbin, wpos, toff=None):
    """Weighted BCE on channel 0 (mean over voxels), softmax cross-entropy
    over channels 1..nbin at voxels with tbin >= 0 (mean over those) and, with
    toff [..., 4] = (dx/8, dy/8, dz/8, weight), the squared error of channels
    nbin+1..nbin+3 against toff[..., :3] summed over the three and averaged
    over voxels of weight > 0. Returns (bce, ce, off)."""
    z = logits[..., 0].reshape(-1)
    t = torch.as_tensor(np.asarray(tprob), dtype=z.dtype).reshape(-1)
    bce = F.binary_cross_entropy_with_logits(z, t, pos_weight=torch.tensor(wpos, dtype=z.dtype), reduction="mean")
    b = torch.as_tensor(np.asarray(tbin)).reshape(-1).long()
    sel = b >= 0
    if int(sel.sum()) > 0:
        L = logits.reshape(-1, HEAD)[sel][:, 1:1 + nbin]
        ce = F.cross_entropy(L, b[sel], reduction="mean")
    else:
        ce = torch.zeros((), dtype=z.dtype)
    off = torch.zeros((), dtype=z.dtype)
    if toff is not None:
        to = torch.as_tensor(np.asarray(toff), dtype=z.dtype).reshape(-1, 4)
        w = to[:, 3] > 0
        if int(w.sum()) > 0:
            O = logits.reshape(-1, HEAD)[w][:, 1 + nbin:4 + nbin]
            off = ((O - to[w][:, :3]) ** 2).sum(1).mean()
    return bce, ce, off


def gradients(vol, theta, tprob, tbin, nbin, wpos, bf16=True, aux=None, toff=None):
    """(bce, ce, off, flat gradient of bce + ce + off) in FP64."""
    P = unpack(theta)
    for w, b in P:
        w.requires_grad_(True)
        b.requires_grad_(True)
    logits, _ = forward(vol, theta, bf16=bf16, params=P, aux=aux)
    bce, ce, off = loss(logits, tprob, tbin, nbin, wpos, toff)
    (bce + ce + off).backward()
    g = []
    for (w, b), (ci, co, t, _) in zip(P, SPEC):
        # torch [co][ci][kz][ky][kx] -> W[t][co][ci]
        g.append(w.grad.reshape(co, ci, t).permute(2, 0, 1).reshape(-1))
        g.append(b.grad.reshape(-1))
    return float(bce), float(ce), float(off), torch.cat(g).detach().numpy()


def conv_direct(x, W, b, relu):
    """Brute-force 3x3x3 convolution of channels-last x [D][H][W][Cin] with
    W[t][co][ci] (zero padding): pins the tap convention."""
    D, H, Wd, Cin = x.shape
    Cout = W.shape[1]
    y = np.zeros((D, H, Wd, Cout))
    for z in range(D):
        for yy in range(H):
            for xx in range(Wd):
                acc = np.array(b, dtype=np.float64)
                for t in range(27):
                    dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < Wd:
                        acc += W[t] @ x[zz, y2, x2]
                y[z, yy, xx] = np.maximum(acc, 0) if relu else acc
    return y
