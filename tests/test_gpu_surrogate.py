"""GPU: the BF16-MFMA surrogate kernels against the FP64 layer oracle.

Tolerances: a layer whose output is stored in bfloat16 may differ from the rounded FP64 value by one
bfloat16 step (relative 2^-7) where FP32 and FP64 sums round to different neighbours; FP32 weight
gradients within 1e-4 relative of FP64; the whole network (eight convolutions, activations and
activation gradients stored in bfloat16) within 2e-2 (outputs) and 6e-2 (gradients) relative L2."""
import numpy as np
import pytest

from beamz_amd import _native as N
from beamz_amd.surrogate import Surrogate
from oracle import surrogate_ref as S

pytestmark = pytest.mark.gpu


def to16(torch, a):
    """FP64 array (already bfloat16-exact) -> int16 device tensor of its bfloat16 bit patterns."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    return torch.from_numpy((f.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)).cuda()


def from16(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def test_conv_layers_match_oracle(hip):
    torch = hip
    rng = np.random.default_rng(0)
    B, H, W, Ci, Co = 2, 8, 12, 64, 96
    x = S.bf16(rng.standard_normal((B, H, W, Ci)))
    w = S.bf16(rng.standard_normal((Co, 3, 3, Ci)) * 0.05)
    b = rng.standard_normal(Co).astype(np.float32).astype(np.float64)
    w32 = torch.tensor(w, dtype=torch.float32).cuda()
    w16 = torch.empty((Co, 3, 3, Ci), dtype=torch.int16).cuda()
    wt16 = torch.empty((Ci, 3, 3, Co), dtype=torch.int16).cuda()
    st = torch.cuda.current_stream().cuda_stream
    N.call("bz_nn_weights", w32.data_ptr(), w16.data_ptr(), wt16.data_ptr(), Co, Ci, st)
    xd = to16(torch, x)
    bd = torch.tensor(b, dtype=torch.float32).cuda()
    y = torch.empty((B, H, W, Co), dtype=torch.int16).cuda()
    N.call("bz_nn_conv3x3", xd.data_ptr(), w16.data_ptr(), bd.data_ptr(), y.data_ptr(), B, H, W, Ci, Co, 1, st)
    ref = np.maximum(S.conv3x3(x, w, b), 0.0)
    got = from16(y)
    assert np.all(np.abs(got - S.bf16(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    # backward data and weights
    gz = S.bf16(rng.standard_normal((B, H, W, Co)))
    gzd = to16(torch, gz)
    gx = torch.empty((B, H, W, Ci), dtype=torch.int16).cuda()
    N.call("bz_nn_conv3x3_dgrad", gzd.data_ptr(), wt16.data_ptr(), gx.data_ptr(), B, H, W, Ci, Co, st)
    dx, dw, db = S.conv3x3_backward(x, w, gz)
    assert np.all(np.abs(from16(gx) - S.bf16(dx)) <= 2.0 ** -7 * np.abs(dx) + 1e-5)
    gw = torch.zeros((Co, 3, 3, Ci), dtype=torch.float32).cuda()
    gb = torch.zeros(Co, dtype=torch.float32).cuda()
    N.call("bz_nn_conv3x3_wgrad", xd.data_ptr(), gzd.data_ptr(), gw.data_ptr(), gb.data_ptr(), B, H, W, Ci, Co, 3, st)
    assert rel_l2(gw.cpu().numpy(), dw) < 1e-4
    assert rel_l2(gb.cpu().numpy(), db) < 1e-5


def test_network_forward_backward_match_oracle(hip):
    torch = hip
    net = Surrogate(seed=3)
    P = {k: v.cpu().numpy().astype(np.float64) for k, v in net.params.items()}
    rng = np.random.default_rng(5)
    x = rng.random((2, 16, 16))
    out = net.forward(torch.tensor(x, dtype=torch.float32).cuda())
    ref, cache = S.forward(P, x.astype(np.float32).astype(np.float64), round_bf16=True)
    assert rel_l2(out.cpu().numpy(), ref) < 2e-2
    gout = rng.standard_normal((2, 8))
    gx = net.backward(torch.tensor(gout, dtype=torch.float32).cuda(), input_grad=True)
    G = S.backward(P, x, cache, gout)
    for k in ("fw", "fb", "w31", "w30", "w21", "w10", "w01", "w00", "b20", "b00"):
        assert rel_l2(net.grads[k].cpu().numpy(), G[k]) < 6e-2, k
    assert rel_l2(gx.cpu().numpy(), G["x"]) < 6e-2


def test_training_reduces_loss(hip):
    torch = hip
    net = Surrogate(seed=1)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = (torch.rand((64, 32, 32), generator=g, device="cuda") > 0.5).float()
    q = x.reshape(64, 2, 16, 2, 16).mean(dim=(2, 4)).reshape(64, 4)
    target = torch.cat([q, 1.0 - q], dim=1)
    first = float(net.train_step(x, target, lr=2e-3))
    for _ in range(60):
        last = float(net.train_step(x, target, lr=2e-3))
    assert np.isfinite(last) and last < 0.2 * first, (first, last)
