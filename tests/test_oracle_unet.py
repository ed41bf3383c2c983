"""Pins the FP64 U-Net oracle (oracle/unet.c; build-defined, no reference
counterpart) against torch float64 autograd running the same network, and
against central finite differences.  A transposed kernel, a wrong tap
offset, a dropped skip connection or a mis-ordered concat fails here."""
import numpy as np
import pytest

from oracle import pyoracle as O

torch = pytest.importorskip("torch")
F = torch.nn.functional


def torch_unet(params, x):
    """The same network in torch (NCHW), float64."""
    ws, bs, off = [], [], 0
    for cin, cout, _ in O.UNET_LAYERS:
        w = params[off:off + cout * 9 * cin].view(cout, 3, 3, cin).permute(0, 3, 1, 2)
        off += cout * 9 * cin
        ws.append(w)
        bs.append(params[off:off + cout])
        off += cout
    relu = [r for _, _, r in O.UNET_LAYERS]

    def conv(X, l):
        y = F.conv2d(X, ws[l], bs[l], padding=1)
        return torch.relu(y) if relu[l] else y

    def up(X):
        return F.interpolate(X, scale_factor=2, mode="nearest")

    X = x.permute(0, 3, 1, 2)
    a2 = conv(conv(X, 0), 1)
    a5 = conv(conv(F.avg_pool2d(a2, 2), 2), 3)
    a8 = conv(conv(F.avg_pool2d(a5, 2), 4), 5)
    a11 = conv(conv(F.avg_pool2d(a8, 2), 6), 7)
    a14 = conv(conv(torch.cat([up(a11), a8], 1), 8), 9)
    a17 = conv(conv(torch.cat([up(a14), a5], 1), 10), 11)
    a20 = conv(conv(torch.cat([up(a17), a2], 1), 12), 13)
    return conv(a20, 14)[:, 0]


@pytest.fixture(scope="module")
def case():
    rng = np.random.default_rng(5)
    params = O.unet_init(3)
    x = rng.standard_normal((2, 16, 24, 8))
    x[..., 4:] = 0
    dout = rng.standard_normal((2, 16, 24))
    return params, x, dout


def test_forward_matches_torch(case):
    params, x, _ = case
    got = O.unet_forward(params, x)
    want = torch_unet(torch.tensor(params), torch.tensor(x)).numpy()
    assert np.abs(want).max() > 1e-3
    assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max())


def test_gradient_matches_torch_autograd(case):
    params, x, dout = case
    got = O.unet_backward(params, x, dout)
    p = torch.tensor(params, requires_grad=True)
    y = torch_unet(p, torch.tensor(x))
    (y * torch.tensor(dout)).sum().backward()
    want = p.grad.numpy()
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_gradient_central_differences(case):
    params, x, dout = case
    g = O.unet_backward(params, x, dout)
    rng = np.random.default_rng(9)
    # sample parameters from every layer's weights and biases
    idx, off = [], 0
    for cin, cout, _ in O.UNET_LAYERS:
        nw = cout * 9 * cin
        rows = 1 if cout == 16 and cin == 16 and off + nw + cout == params.size else cout
        idx += list(off + rng.integers(0, rows * 9 * cin, 2))
        idx.append(off + nw + rng.integers(0, rows))
        off += nw + cout
    for i in idx:
        e = 1e-6 * max(1.0, abs(params[i]))
        pp, pm = params.copy(), params.copy()
        pp[i] += e
        pm[i] -= e
        fd = ((O.unet_forward(pp, x) - O.unet_forward(pm, x)) * dout).sum() / (2 * e)
        assert abs(fd - g[i]) < 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_emulation_is_close_not_equal(case):
    params, x, dout = case
    a = O.unet_forward(params, x)
    b = O.unet_forward(params, x, emulate_bf16=True)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-5 < rel < 5e-2
    ga = O.unet_backward(params, x, dout)
    gb = O.unet_backward(params, x, dout, emulate_bf16=True)
    cos = ga @ gb / np.linalg.norm(ga) / np.linalg.norm(gb)
    assert cos > 0.995
