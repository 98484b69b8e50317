"""The U-Net's hand-written kernels against a PyTorch FP32 reference of the same
network: BF16 weights, activations rounded to BF16 at the points the GPU
stores them (straight-through in the backward pass), FP32 accumulation."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

import serac_amd as sx

pytestmark = pytest.mark.gpu


class RoundBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).float()

    @staticmethod
    def backward(ctx, g):
        return g


R = RoundBF16.apply


def torch_params(net, p):
    """Split the flat parameter vector into torch conv weights [co][ci][kz][ky][kx] and biases."""
    out, off = [], 0
    for ca, cb, co in sx.UNet.LAYERS:
        ci = ca + cb
        w = p[off:off + 27 * ci * co].reshape(3, 3, 3, ci, co)
        off += 27 * ci * co
        b = p[off:off + co]
        off += co
        W = torch.tensor(np.ascontiguousarray(w.transpose(4, 3, 0, 1, 2)), dtype=torch.float32, requires_grad=True)
        B = torch.tensor(b, dtype=torch.float32, requires_grad=True)
        out.append((W, B))
    return out


def to_torch(x, dims, C):
    nx, ny, nz = dims
    return torch.tensor(np.asarray(x, dtype=np.float32).reshape(nz, ny, nx, C).transpose(3, 0, 1, 2)[None])


def grad_flat(params):
    g = []
    for W, B in params:
        g.append(W.grad.detach().numpy().transpose(2, 3, 4, 1, 0).ravel())
        g.append(B.grad.detach().numpy())
    return np.concatenate(g)


def reference(params, x, dims, hidden_relu=True):
    def conv(x, k, stride=1, relu=True):
        W, B = params[k]
        y = F.conv3d(x, R(W), B, stride=stride, padding=1)
        if not relu:
            return y
        return R(torch.relu(y)) if hidden_relu else R(y)

    def up(x, like):
        return R(F.interpolate(x, size=like.shape[2:], mode="trilinear", align_corners=True))

    x = R(x)
    a0 = conv(x, 0)
    b0 = conv(a0, 1)
    a1 = conv(b0, 2, 2)
    b1 = conv(a1, 3)
    a2 = conv(b1, 4, 2)
    b2 = conv(a2, 5)
    a3 = conv(b2, 6, 2)
    b3 = conv(a3, 7)
    c2 = conv(torch.cat([up(b3, b2), b2], 1), 8)
    c1 = conv(torch.cat([up(c2, b1), b1], 1), 9)
    c0 = conv(torch.cat([up(c1, b0), b0], 1), 10)
    return conv(c0, 11, relu=False)


# below 32768 output nodes a convolution runs one lane per (node, channel);
# the 41 x 33 x 33 grid also exercises the one-thread-per-node kernel on its
# finest level
@pytest.mark.parametrize("dims,act", [((17, 17, 17), "relu"), ((17, 9, 33), "relu"), ((41, 33, 33), "relu"),
                                      ((17, 17, 17), "identity")])
def test_unet_forward_backward_match_torch(gpu_lib, dims, act):
    net = sx.UNet(*dims, seed=3)
    if act == "identity":
        net.set_activation(sx.capi.SX_UNET_IDENTITY, seed=3)
    rng = np.random.default_rng(5)
    # give the biases some value so that they are exercised
    p = net.params()
    off = 0
    for ca, cb, co in sx.UNet.LAYERS:
        off += 27 * (ca + cb) * co
        p[off:off + co] = rng.uniform(-0.05, 0.05, co)
        off += co
    net.set_params(p)
    x = rng.standard_normal((net.nvox, 3)).astype(np.float32)
    t = rng.standard_normal((net.nvox, 3)).astype(np.float32)
    y = net.forward(x)
    params = torch_params(net, net.params())
    yt = reference(params, to_torch(x, dims, 3), dims, hidden_relu=(act == "relu"))
    y_ref = yt.detach().numpy()[0].transpose(1, 2, 3, 0).reshape(-1, 3)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < 1e-2
    assert np.abs(y - y_ref).max() < 5e-2 * np.abs(y_ref).max()
    # gradients of the relative squared error
    loss = net.backward(x, t)
    tt = to_torch(t, dims, 3)
    lt = ((yt - tt) ** 2).sum() / (tt ** 2).sum()
    lt.backward()
    assert abs(loss - lt.item()) < 1e-2 * lt.item()
    g, g_ref = net.grads(), grad_flat(params)
    # BF16 activations round differently where the FP32 sums differ in order:
    # a few percent on the coarse levels (few nodes), far below the O(1) error
    # of a wrong tap, stride or transpose
    assert np.linalg.norm(g - g_ref) <= 2e-2 * np.linalg.norm(g_ref)
    off = 0
    for k, (ca, cb, co) in enumerate(sx.UNet.LAYERS):
        n = 27 * (ca + cb) * co + co
        a, b = g[off:off + n], g_ref[off:off + n]
        assert np.linalg.norm(a - b) <= 8e-2 * np.linalg.norm(b), f"layer {k}"
        off += n


@pytest.mark.parametrize("dims", [(17, 17, 17), (33, 17, 9), (65, 65, 65)])
def test_mfma_forward_matches_vector_forward(gpu_lib, dims):
    # the matrix-core convolutions against the vector-unit ones: the same BF16
    # operands, FP32 sums in another order
    net = sx.UNet(*dims, seed=6)
    rng = np.random.default_rng(8)
    p = net.params()
    off = 0
    for ca, cb, co in sx.UNet.LAYERS:
        off += 27 * (ca + cb) * co
        p[off:off + co] = rng.uniform(-0.05, 0.05, co)
        off += co
    net.set_params(p)
    x = rng.standard_normal((net.nvox, 3)).astype(np.float32)
    net.set_mfma(False)
    y_vec = net.forward(x)
    net.set_mfma(True)
    y_mfma = net.forward(x)
    assert np.linalg.norm(y_mfma - y_vec) <= 1e-2 * np.linalg.norm(y_vec)


def test_identity_initialisation_scales(gpu_lib):
    # identity activation: weights ~ N(0, 1 / fan_in), output layer 1e-2 of that
    net = sx.UNet(17, 17, 17, seed=4)
    net.set_activation(sx.capi.SX_UNET_IDENTITY, seed=4)
    p = net.params()
    off = 0
    for k, (ca, cb, co) in enumerate(sx.UNet.LAYERS):
        nw = 27 * (ca + cb) * co
        w = p[off:off + nw]
        expect = np.sqrt(1.0 / (27 * (ca + cb))) * (1e-2 if k == 11 else 1.0)
        assert abs(w.std() / expect - 1.0) < 0.15, k
        assert np.all(p[off + nw:off + nw + co] == 0.0)
        off += nw + co


def test_unet_adam_reduces_loss_on_a_fixed_pair(gpu_lib):
    net = sx.UNet(17, 17, 17, seed=1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((net.nvox, 3)).astype(np.float32)
    t = 0.1 * x  # a linear target the network can fit
    losses = []
    for _ in range(30):
        losses.append(net.backward(x, t))
        net.adam(1e-3)
    assert losses[-1] < 0.5 * losses[0]
    assert np.abs(net.grads()).max() == 0.0  # Adam cleared the accumulated gradients
