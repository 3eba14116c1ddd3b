"""CPU reference of the learned preconditioner's U-Net (torch, float64).

Test infrastructure only: the product runs the hand-written HIP kernels of
continuum_liquid_amd/csrc/unet.hip.  The reference reads the same parameter
vector (layout in include/cliquid.h, clq_net_*) and computes the same map:

    s     = sqrt(mean over interior cells with mask > 0 of rhs^2),
    k     = s (N h)^2 / 1024 with h = dx and N the largest grid dimension
    input = [bf16(rhs / s), mask > 0, bf16(prev / k), 0 x 13] on the grid
            rounded up to multiples of 4 (zero padding)
    c1..c10 = 3x3x3 conv, zero padding, bias, ReLU, then rounded to bf16;
            average pooling 2x2x2 and nearest upsampling + concat as listed
    out   = prev + k * (head . a10 + head_bias), cropped to the grid
            (prev = 0 when absent)

With `emulate=True` weights and every stored activation are rounded to bf16
(round to nearest even), as on the device; the rounding passes gradients
straight through, so autograd gives the gradient of the function the device
evaluates (up to the FP32 accumulation and the bf16 rounding of dL/dY that the
device's backward kernels apply).
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = ((16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16))


def param_count() -> int:
    return sum(27 * ci * co + co for ci, co in LAYERS) + 17


def unpack(params) -> tuple[list, torch.Tensor, torch.Tensor]:
    """Parameter vector -> [(weight [co, ci, 3, 3, 3], bias [co])], head w, head b.

    The vector stores W[tap][ci][co] with tap = 9 (dx+1) + 3 (dy+1) + (dz+1);
    torch's conv3d weight is [co][ci][kx][ky][kz] with k = d + 1."""
    p = torch.as_tensor(np.asarray(params, np.float64))
    out, off = [], 0
    for ci, co in LAYERS:
        w = p[off:off + 27 * ci * co].reshape(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2)
        off += 27 * ci * co
        out.append((w, p[off:off + co]))
        off += co
    return out, p[off:off + 16], p[off + 16]


def bf16(t: torch.Tensor) -> torch.Tensor:
    """Round to bfloat16 (nearest even) with a straight-through gradient."""
    r = t.detach().to(torch.bfloat16).to(t.dtype)
    return t + (r - t.detach())


def padded_shape(shape):
    return tuple((n + 3) // 4 * 4 for n in shape)


def input_scale(rhs: np.ndarray, mask: np.ndarray, dx: float) -> tuple:
    """(1/s, k, 1/k, count) as the device computes them (FP64 sums and
    products, FP32 results; h is dx rounded to FP32)."""
    inner = np.zeros(rhs.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    sel = inner & (mask > 0)
    n = int(sel.sum())
    rms = np.sqrt(np.sum(rhs[sel].astype(np.float64) ** 2) / n) if n else 0.0
    if rms == 0.0:
        return np.float32(0), np.float32(0), np.float32(0), n
    extent = float(np.float32(dx)) * max(rhs.shape)
    len2 = extent * extent / 1024.0
    return np.float32(1.0 / rms), np.float32(rms * len2), np.float32(1.0 / (rms * len2)), n


def forward(params, rhs, mask, prev=None, dx: float = 1.0, *, emulate: bool = True, weights=None):
    """Network output on the grid (float64 tensor) and the tensors that carry
    gradients: `weights` (list of leaf tensors) when given are used in place
    of the unpacked parameters."""
    shape = rhs.shape
    inv_s, out_scale, inv_out, _ = input_scale(rhs, mask, dx)
    rnd = bf16 if emulate else (lambda t: t)
    layers, hw, hb = weights if weights is not None else unpack(params)
    X, Y, Z = padded_shape(shape)
    x = torch.zeros((1, 16, X, Y, Z), dtype=torch.float64)
    c0 = torch.as_tensor((rhs.astype(np.float32) * inv_s).astype(np.float64))
    x[0, 0, :shape[0], :shape[1], :shape[2]] = rnd(c0)
    x[0, 1, :shape[0], :shape[1], :shape[2]] = torch.as_tensor((mask > 0).astype(np.float64))
    if prev is not None:
        c2 = torch.as_tensor((prev.astype(np.float32) * inv_out).astype(np.float64))
        x[0, 2, :shape[0], :shape[1], :shape[2]] = rnd(c2)

    def conv(t, i):
        w, b = layers[i]
        return rnd(F.relu(F.conv3d(t, rnd(w), b, padding=1)))

    def pool(t):
        return rnd(F.avg_pool3d(t, 2))

    def up(t):
        return F.interpolate(t, scale_factor=2, mode="nearest")

    a1 = conv(x, 0)
    s0 = conv(a1, 1)
    a3 = conv(pool(s0), 2)
    s1 = conv(a3, 3)
    a5 = conv(pool(s1), 4)
    a6 = conv(a5, 5)
    a7 = conv(torch.cat([up(a6), s1], 1), 6)
    a8 = conv(a7, 7)
    a9 = conv(torch.cat([up(a8), s0], 1), 8)
    a10 = conv(a9, 9)
    pred = torch.einsum("c,cxyz->xyz", hw, a10[0]) + hb
    out = float(out_scale) * pred[:shape[0], :shape[1], :shape[2]]
    if prev is not None:
        out = out + torch.as_tensor(prev.astype(np.float64))
    return out


def gradients(params, rhs, mask, prev, grad_out, dx: float = 1.0, *, emulate: bool = True) -> np.ndarray:
    """Gradient of sum(grad_out * out) with respect to the parameter vector."""
    layers, hw, hb = unpack(params)
    leaves = []
    for w, b in layers:
        w = w.clone().requires_grad_(True)
        b = b.clone().requires_grad_(True)
        leaves.append((w, b))
    hw = hw.clone().requires_grad_(True)
    hb = hb.clone().requires_grad_(True)
    out = forward(params, rhs, mask, prev, dx, emulate=emulate, weights=(leaves, hw, hb))
    (out * torch.as_tensor(grad_out.astype(np.float64))).sum().backward()
    parts = []
    for w, b in leaves:
        parts.append(w.grad.permute(2, 3, 4, 1, 0).reshape(-1))
        parts.append(b.grad)
    parts += [hw.grad, hb.grad.reshape(1)]
    return torch.cat(parts).numpy()


def layer_slices():
    """Slices of the parameter vector: (name, slice) per weight/bias block."""
    out, off = [], 0
    for i, (ci, co) in enumerate(LAYERS):
        out.append((f"c{i + 1}.weight", slice(off, off + 27 * ci * co)))
        off += 27 * ci * co
        out.append((f"c{i + 1}.bias", slice(off, off + co)))
        off += co
    out.append(("head.weight", slice(off, off + 16)))
    out.append(("head.bias", slice(off + 16, off + 17)))
    return out
