"""Learned preconditioner network (stratiflow_amd.Net, sf_net_*).

CPU: the numpy restatement (oracle/convnet.py) against FP64 finite differences
and against closed forms (identity taps, linearity in the last layer).
GPU: the MFMA forward pass and parameter gradients against the restatement
with the same BF16 roundings; training lowers the loss; the preconditioner is
the identity until trained.
"""
import numpy as np
import pytest

CH = [32, 64, 64, 64, 32]


def _rand_params(rng, channels, creal_in, creal_out, scale=1.0, last_zero=False):
    from oracle import convnet
    layers = []
    for l in range(len(channels) - 1):
        ci, co = channels[l], channels[l + 1]
        rin = creal_in if l == 0 else ci
        rout = creal_out if l == len(channels) - 2 else co
        W = np.zeros((5, co, ci))
        W[:, :rout, :rin] = rng.standard_normal((5, rout, rin)) * scale * np.sqrt(2.0 / (5 * rin))
        if last_zero and l == len(channels) - 2:
            W[:] = 0
        b = np.zeros(co)
        b[:rout] = 0.1 * rng.standard_normal(rout)
        layers.append((W, b))
    return convnet.pack(layers)


@pytest.mark.parametrize("dil", [None, [1, 2, 3]])
def test_restatement_gradient_matches_finite_differences(dil):
    from oracle import convnet
    rng = np.random.default_rng(0)
    ch = [4, 6, 5, 3]
    p = _rand_params(rng, ch, 4, 3)
    x = rng.standard_normal((2, 11, 4))
    tgt = rng.standard_normal((2, 11, 3))
    loss, g = convnet.loss_and_grad(p, ch, x, tgt, 3, bf16=False, dil=dil)
    eps = 1e-6
    for k in rng.choice(p.size, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (convnet.loss_and_grad(pp, ch, x, tgt, 3, bf16=False, dil=dil)[0]
              - convnet.loss_and_grad(pm, ch, x, tgt, 3, bf16=False, dil=dil)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_restatement_closed_forms():
    """A centre-tap identity layer copies its input; profile ends see zero padding."""
    from oracle import convnet
    x = np.arange(12, dtype=float).reshape(1, 6, 2)
    W = np.zeros((5, 2, 2))
    W[2] = np.eye(2)
    assert np.allclose(convnet.conv(x, W, None), x)
    W = np.zeros((5, 2, 2))
    W[4] = np.eye(2)  # out[p] = in[p + 2]
    out = convnet.conv(x, W, None)
    assert np.allclose(out[0, :4], x[0, 2:]) and np.allclose(out[0, 4:], 0)
    out = convnet.conv(x, W, None, 2)  # dilation 2: out[p] = in[p + 4]
    assert np.allclose(out[0, :2], x[0, 4:]) and np.allclose(out[0, 2:], 0)
    W = np.zeros((5, 2, 2))
    W[0] = np.eye(2)  # out[p] = in[p - 2 d]
    out = convnet.conv(x, W, None, 2)
    assert np.allclose(out[0, 4:], x[0, :2]) and np.allclose(out[0, :4], 0)
    assert convnet.to_bf16(np.float32(1.0 + 2 ** -9)) == np.float32(1.0)
    assert convnet.to_bf16(np.float32(1.0 + 3 * 2 ** -9)) == np.float32(1.0 + 2 ** -7)


def _net(N1=48, N3=64, prec=4, seed=3):
    import stratiflow_amd as sf
    s = sf.IMEXRK(N1, 1, N3, sf.TWO_D, 2 * np.pi / 0.44, 1.0, 10.0, 1000.0, 0.1, 1.0, precision=prec)
    return s, sf.Net(s, hidden=64, layers=4, seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("dil", [None, [1, 2, 4, 8]])
def test_forward_and_gradient_match_restatement(dil):
    from oracle import convnet
    s, net = _net()
    assert net.channels == CH
    if dil:
        net.set_dilations(dil)
    rng = np.random.default_rng(1)
    p = _rand_params(rng, CH, 8, 6).astype(np.float32)
    net.set_params(p)
    x = np.zeros((5, net.Z, 32), np.float32)
    x[..., :8] = rng.standard_normal((5, net.Z, 8))
    out = net.forward_raw(x)
    ref, _, _ = convnet.forward(p, CH, x, bf16=True, dil=dil)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print("forward rel err", err)
    assert err < 2e-3
    tgt = np.zeros_like(out)
    tgt[..., :6] = rng.standard_normal((5, net.Z, 6))
    loss, g = net.grad_raw(x, tgt)
    lref, gref = convnet.loss_and_grad(p, CH, x, tgt, 6, bf16=True, dil=dil)
    print("loss gpu %.6e ref %.6e" % (loss, lref))
    assert abs(loss - lref) / lref < 1e-3
    eg = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print("gradient rel err", eg)
    assert eg < 2e-2


@pytest.mark.gpu
def test_training_lowers_loss_and_identity_until_trained():
    import stratiflow_amd as sf
    from stratiflow_amd import ic
    s, net = _net()
    z, zf = s.vertical_points()
    pert = ic.perturbation(48, 1, 64, sf.TWO_D, z, zf, 5, amplitude=1e-2)
    for n, a in pert.items():
        s.set(n, a)
    fin = [s.state(n) for n in ("u1", "u2", "u3", "b")]
    outs = [s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN) for n in ("u1", "u2", "u3", "b")]
    net.apply(fin, outs)
    for f, o in zip(fin, outs):
        assert np.array_equal(f.download(), o.download())
    # pairs of a fixed per-field scaling q = (0.5 u1, 2 u3, b): the direction
    # target q/RMS(q) - w/RMS(w) is a per-channel scaling of the input
    scale = {"u1": 0.5, "u2": 1.0, "u3": 2.0, "b": 1.0}
    for k in range(6):
        w = [s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN) for n in ("u1", "u2", "u3", "b")]
        q = [s.field(sf.MODAL, sf.DIRICHLET if n == "u3" else sf.NEUMANN) for n in ("u1", "u2", "u3", "b")]
        pk = ic.perturbation(48, 1, 64, sf.TWO_D, z, zf, 100 + k, amplitude=1.0)
        for i, n in enumerate(("u1", "u2", "u3", "b")):
            if n in pk:
                w[i].upload(pk[n])
                q[i].upload(scale[n] * pk[n])
        net.add_pair(w, q)
    assert net.samples() == 6 * 16
    l0, l1 = net.train(epochs=60, batch=32, lr=2e-3)
    print("loss first epoch %.4e last %.4e" % (l0, l1))
    assert l1 < 0.1 * l0
