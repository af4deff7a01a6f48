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


def test_batched_weight_copies_equal_per_layer_copies(hip):
    """bz_nn_weights_n (every layer in one launch) writes the bits bz_nn_weights writes layer by layer, also for
    ragged layer sizes and a layer without the transposed copy."""
    import ctypes
    torch = hip
    rng = np.random.default_rng(7)
    shapes = [(32, 1), (64, 32), (96, 40), (8, 8), (256, 256)]
    st = torch.cuda.current_stream().cuda_stream
    ws = [torch.tensor(rng.standard_normal((co, 3, 3, ci)), dtype=torch.float32).cuda() for co, ci in shapes]
    one = [(torch.empty((co, 3, 3, ci), dtype=torch.int16).cuda(), torch.empty((ci, 3, 3, co), dtype=torch.int16).cuda())
           for co, ci in shapes]
    many = [(torch.full((co, 3, 3, ci), 7, dtype=torch.int16).cuda(),
             torch.full((ci, 3, 3, co), 7, dtype=torch.int16).cuda()) for co, ci in shapes]
    for (co, ci), w, (a, b) in zip(shapes, ws, one):
        N.call("bz_nn_weights", w.data_ptr(), a.data_ptr(), b.data_ptr() if co != 8 else None, co, ci, st)
    n = len(shapes)
    P, I = ctypes.c_void_p * n, ctypes.c_int * n
    arrs = (P(*[w.data_ptr() for w in ws]), P(*[a.data_ptr() for a, _ in many]),
            P(*[b.data_ptr() if co != 8 else None for (co, _), (_, b) in zip(shapes, many)]),
            I(*[co for co, _ in shapes]), I(*[ci for _, ci in shapes]))
    N.call("bz_nn_weights_n", *[ctypes.addressof(x) for x in arrs], n, st)
    torch.cuda.synchronize()
    for (co, ci), (a, b), (c, d) in zip(shapes, one, many):
        assert torch.equal(a, c), (co, ci)
        if co != 8:
            assert torch.equal(b, d), (co, ci)
        else:
            assert bool((d == 7).all())


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
    N.call("bz_nn_conv3x3_wgrad", xd.data_ptr(), gzd.data_ptr(), gw.data_ptr(), gb.data_ptr(), B, H, W, Ci, Co, 3,
           None, 0, st)
    assert rel_l2(gw.cpu().numpy(), dw) < 1e-4
    assert rel_l2(gb.cpu().numpy(), db) < 1e-5
    # ordered reductions through a workspace: the same sums, and the same bits on every call
    need = N.lib().bz_nn_wgrad_work(B, H, W, Ci, Co, 3)
    work = torch.empty(need, dtype=torch.float32).cuda()
    with pytest.raises(ValueError, match="workspace"):
        N.call("bz_nn_conv3x3_wgrad", xd.data_ptr(), gzd.data_ptr(), gw.data_ptr(), gb.data_ptr(), B, H, W, Ci, Co, 3,
               work.data_ptr(), need - 1, st)
    runs = []
    for _ in range(2):
        gw.zero_(); gb.zero_()
        N.call("bz_nn_conv3x3_wgrad", xd.data_ptr(), gzd.data_ptr(), gw.data_ptr(), gb.data_ptr(), B, H, W, Ci, Co,
               3, work.data_ptr(), need, st)
        runs.append((gw.cpu().numpy(), gb.cpu().numpy()))
    assert rel_l2(runs[0][0], dw) < 1e-4 and rel_l2(runs[0][1], db) < 1e-5
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("B,H,W,Ci,Co", [(2, 8, 16, 32, 96), (1, 2, 128, 32, 64), (3, 4, 32, 32, 32),
                                         (2, 8, 16, 64, 32), (2, 5, 128, 32, 32), (3, 4, 64, 32, 64),
                                         (2, 3, 64, 64, 64), (1, 9, 64, 32, 32), (3, 1, 128, 32, 32),
                                         (1, 17, 64, 64, 32)])
def test_halo_convolutions_match_oracle(hip, B, H, W, Ci, Co):
    """Forward and backward-data convolutions on the halo-staged kernel's shapes (32-channel inputs on 64-pixel
    tiles of whole rows or row segments, an output count that is not a multiple of 64; the fourth case takes
    it for backward-data only) and on the row march's (rows of 64 or 128 pixels: every channel-block shape,
    bands spanning images, single-row images, one image over two bands), within one bfloat16 step of the
    rounded FP64 oracle."""
    torch = hip
    rng = np.random.default_rng(B * 1000 + W)
    x = S.bf16(rng.standard_normal((B, H, W, Ci)))
    w = S.bf16(rng.standard_normal((Co, 3, 3, Ci)) * 0.05)
    b = rng.standard_normal(Co).astype(np.float32).astype(np.float64)
    w32 = torch.tensor(w, dtype=torch.float32).cuda()
    w16 = torch.empty((Co, 3, 3, Ci), dtype=torch.int16).cuda()
    wt16 = torch.empty((Ci, 3, 3, Co), dtype=torch.int16).cuda()
    st = torch.cuda.current_stream().cuda_stream
    N.call("bz_nn_weights", w32.data_ptr(), w16.data_ptr(), wt16.data_ptr(), Co, Ci, st)
    y = torch.empty((B, H, W, Co), dtype=torch.int16).cuda()
    # device inputs stay referenced until the kernels have run (a temporary's memory can be handed to the
    # next allocation, whose upload is queued before the kernel)
    xd, bd = to16(torch, x), torch.tensor(b, dtype=torch.float32).cuda()
    N.call("bz_nn_conv3x3", xd.data_ptr(), w16.data_ptr(), bd.data_ptr(), y.data_ptr(), B, H, W, Ci, Co, 1, st)
    ref = np.maximum(S.conv3x3(x, w, b), 0.0)
    assert np.all(np.abs(from16(y) - S.bf16(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    gz = S.bf16(rng.standard_normal((B, H, W, Co)))
    gx = torch.empty((B, H, W, Ci), dtype=torch.int16).cuda()
    gzd = to16(torch, gz)
    N.call("bz_nn_conv3x3_dgrad", gzd.data_ptr(), wt16.data_ptr(), gx.data_ptr(), B, H, W, Ci, Co, st)
    dx, _, _ = S.conv3x3_backward(x, w, gz)
    assert np.all(np.abs(from16(gx) - S.bf16(dx)) <= 2.0 ** -7 * np.abs(dx) + 1e-5)


def test_first_layer_pooling_and_mask_match_oracle(hip):
    """The one-input-channel first convolution (forward, weight and bias gradients), 2x2 average pooling and
    its backward, and the ReLU mask, on shapes whose channel counts take the 8-wide kernels (32, 64) and on
    one that takes the scalar ones (12 channels for pooling, an odd element count for the mask)."""
    torch = hip
    rng = np.random.default_rng(11)
    st = torch.cuda.current_stream().cuda_stream
    B, H, W, Co = 3, 10, 14, 32
    x = S.bf16(rng.random((B, H, W, 1)))
    w = S.bf16(rng.standard_normal((Co, 3, 3, 1)) * 0.3)
    b = rng.standard_normal(Co).astype(np.float32).astype(np.float64) * 0.1
    xd = to16(torch, x)
    wd = torch.tensor(w.reshape(Co, 9), dtype=torch.float32).cuda()
    bd = torch.tensor(b, dtype=torch.float32).cuda()
    y = torch.empty((B, H, W, Co), dtype=torch.int16).cuda()
    N.call("bz_nn_conv_first", xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), y.data_ptr(), B, H, W, Co, st)
    ref = np.maximum(S.conv3x3(x, w, b), 0.0)
    assert np.all(np.abs(from16(y) - S.bf16(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    gz = S.bf16(rng.standard_normal((B, H, W, Co)))
    gw = torch.zeros((Co, 9), dtype=torch.float32).cuda()
    gb = torch.zeros(Co, dtype=torch.float32).cuda()
    _, dw, db = S.conv3x3_backward(x, w, gz)
    need = N.lib().bz_nn_first_work(B, H, W, Co)
    work = torch.empty(need, dtype=torch.float32).cuda()
    for wk, nwk in ((None, 0), (work.data_ptr(), need)):           # atomics, then ordered partial sums
        gw.zero_(); gb.zero_()
        gzd = to16(torch, gz)
        N.call("bz_nn_conv_first_bwd", xd.data_ptr(), gzd.data_ptr(), wd.data_ptr(), gw.data_ptr(),
               gb.data_ptr(), None, B, H, W, Co, wk, nwk, st)
        assert rel_l2(gw.cpu().numpy(), dw.reshape(Co, 9)) < 1e-5
        assert rel_l2(gb.cpu().numpy(), db) < 1e-5
    for C in (32, 12):                      # input gradient: 16-byte channel loads (C % 8 == 0), then per channel
        wc = S.bf16(rng.standard_normal((C, 3, 3, 1)) * 0.3)
        gzc = S.bf16(rng.standard_normal((B, H, W, C)))
        gx = torch.empty((B, H, W), dtype=torch.float32).cuda()
        wcd = torch.tensor(wc.reshape(C, 9), dtype=torch.float32).cuda()
        gwc = torch.zeros((C, 9), dtype=torch.float32).cuda()
        gbc = torch.zeros(C, dtype=torch.float32).cuda()
        gzcd = to16(torch, gzc)
        N.call("bz_nn_conv_first_bwd", xd.data_ptr(), gzcd.data_ptr(), wcd.data_ptr(), gwc.data_ptr(),
               gbc.data_ptr(), gx.data_ptr(), B, H, W, C, None, 0, st)
        dx, _, _ = S.conv3x3_backward(x, wc, gzc)
        assert rel_l2(gx.cpu().numpy(), dx[..., 0]) < 1e-5, C
    for C in (64, 12):
        a = S.bf16(rng.standard_normal((2, 6, 8, C)))
        p = torch.empty((2, 3, 4, C), dtype=torch.int16).cuda()
        ad = to16(torch, a)
        N.call("bz_nn_avgpool2", ad.data_ptr(), p.data_ptr(), 2, 6, 8, C, st)
        pr = S.avgpool2(a)
        assert np.all(np.abs(from16(p) - S.bf16(pr)) <= 2.0 ** -7 * np.abs(pr) + 1e-7), C
        gy = S.bf16(rng.standard_normal((2, 3, 4, C)))
        gx = torch.empty((2, 6, 8, C), dtype=torch.int16).cuda()
        gyd = to16(torch, gy)
        N.call("bz_nn_avgpool2_bwd", gyd.data_ptr(), gx.data_ptr(), 2, 6, 8, C, st)
        assert np.array_equal(from16(gx), S.avgpool2_backward(gy)), C      # a quarter is exact in bfloat16
    for n in (4096, 4099):
        g = S.bf16(rng.standard_normal(n))
        v = S.bf16(rng.standard_normal(n))
        gd = to16(torch, g)
        vd = to16(torch, v)
        N.call("bz_nn_relu_mask", gd.data_ptr(), vd.data_ptr(), n, st)
        assert np.array_equal(from16(gd), np.where(v > 0, g, 0.0)), n


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


def test_training_is_deterministic(hip):
    """Two networks from the same seed, trained on the same batches, hold the same bits: every gradient
    reduction is added in a fixed order."""
    torch = hip
    g = torch.Generator(device="cuda").manual_seed(1)
    x = (torch.rand((96, 32, 32), generator=g, device="cuda") > 0.5).float()
    target = torch.rand((96, 8), generator=g, device="cuda")
    params = []
    for _ in range(2):
        net = Surrogate(seed=4)
        for q in range(0, 96, 32):
            net.train_step(x[q:q + 32], target[q:q + 32], lr=1e-3)
        params.append({k: v.cpu().numpy() for k, v in net.params.items()})
    for k in params[0]:
        assert np.array_equal(params[0][k], params[1][k]), k


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


def test_surrogate_trained_on_fdtd_data_beats_random_screening(hip):
    """The surrogate, trained on 4096 spectra generated by the batched FDTD, must (a) predict held-out
    spectra better than the constant predictor (the training mean), the method without the network, and
    (b) rank 1024 fresh candidates so that its top-16 picks verify (by FDTD) above a random pick of 16."""
    from beamz_amd.surrogate.data import TransmissionProblem, generate, random_densities
    from beamz_amd.surrogate.optimize import relative_error, screen, train
    prob = TransmissionProblem(batch=256)
    rho, T = generate(prob, 4096, seed=1)
    assert np.all(np.isfinite(T)) and T.shape == (4096, 8)
    net = Surrogate(seed=0)
    hist = train(net, rho, T, epochs=6, batch=128, lr=5e-4)
    rho_v, T_v = generate(prob, 512, seed=2)
    err, _ = relative_error(net, rho_v, T_v)
    err_const = float(np.linalg.norm(T.mean(axis=0) - T_v) / np.linalg.norm(T_v))
    cands = random_densities(1024, seed=3)
    fom_best, fom_rand, _ = screen(net, prob, cands, top=16, seed=4)
    print(f"surrogate: train loss {hist[0]:.3e} -> {hist[-1]:.3e}, held-out rel-L2 {err:.3f} "
          f"(constant predictor {err_const:.3f}), verified FOM top-16 {fom_best.mean():.4f} "
          f"vs random-16 {fom_rand.mean():.4f}")
    assert hist[-1] < hist[0]
    assert err < err_const
    assert fom_best.mean() > fom_rand.mean()


def test_transmission_labels_follow_their_densities(hip):
    """The batched label generator must give each density its own spectrum: permuting the densities (across
    two batch chunks, the second one partly filled) permutes the spectra bit for bit, and a rerun after
    reset reproduces them."""
    from beamz_amd.surrogate.data import TransmissionProblem, random_densities
    prob = TransmissionProblem(batch=4, steps=400)
    rho = random_densities(6, seed=7)
    T1 = prob.transmission(rho)
    perm = [5, 2, 0, 4, 1, 3]
    T2 = prob.transmission(rho[perm])
    assert np.all(np.isfinite(T1)) and T1.std() > 0
    assert np.array_equal(T2, T1[perm])
    assert np.array_equal(prob.transmission(rho), T1)


def test_surrogate_loop_rounds_verify_above_random_and_report_the_error(hip):
    """The surrogate-assisted loop of the surrogate workload, at reduced size: a network trained on 2048 FDTD
    spectra, then 3 rounds of propose 1024 (mutations of the best verified designs climbed along the network's
    input gradient) -> verify the top 16 by FDTD -> fine-tune. Each round's verified picks must beat a random
    pick of 16 fresh designs on average, the best verified figure never falls, and the network's error against
    FDTD on its own picks is reported (the plain FDTD loop and the time-to-FOM comparison are bench.py's
    surrogate workload)."""
    from beamz_amd.surrogate.data import TransmissionProblem, generate, random_densities
    from beamz_amd.surrogate.optimize import SurrogateLoop, train
    prob = TransmissionProblem(batch=256)
    rho, T = generate(prob, 2048, seed=5)
    net = Surrogate(seed=1)
    train(net, rho, T, epochs=4, batch=128, lr=5e-4)
    fom = T.mean(axis=1)
    top = np.argsort(-fom)[:64]
    loop = SurrogateLoop(net, prob, candidates=1024, top=16, seed=6)
    best = loop.run(rho[top], T[top], rounds=3)
    rand = prob.transmission(random_densities(16, seed=8)).mean(axis=1)
    for q, r in enumerate(loop.rounds):
        print(f"round {q}: verified FOM mean {r['verified_fom_mean']:.4f} max {r['verified_fom_max']:.4f}, "
              f"predicted {r['predicted_fom_mean']:.4f}, surrogate rel-L2 vs FDTD {r['rel_error']:.3f}")
    print(f"seed best {fom.max():.4f}, loop best {best:.4f}, random-16 mean {rand.mean():.4f}")
    assert len(loop.rounds) == 3 and all(np.isfinite(r["rel_error"]) for r in loop.rounds)
    assert best >= fom.max()
    bests = [b for _, b in loop.history]
    assert all(b1 >= b0 for b0, b1 in zip(bests, bests[1:]))
    assert all(r["verified_fom_mean"] > rand.mean() for r in loop.rounds)


@pytest.mark.parametrize("B,H,W,Ci,Co", [(2, 5, 128, 32, 32), (4, 10, 64, 32, 64), (2, 3, 64, 64, 64),
                                         (1, 4, 64, 64, 32), (3, 1, 64, 32, 32), (1, 9, 128, 64, 32),
                                         (2, 6, 32, 64, 128), (3, 5, 32, 32, 64), (1, 7, 32, 64, 128),
                                         (4, 9, 16, 64, 64), (1, 16, 16, 128, 64), (5, 1, 16, 64, 64)])
def test_row_march_weight_gradients_match_oracle(hip, B, H, W, Ci, Co):
    """Weight and bias gradients of rows of 16 to 128 pixels (the row-march kernel, with the ordered
    reduction): every channel-block shape and rows per step, bands that span images, a band whose last step
    is short, single-row images (both outer kernel rows skipped) and one image over several bands; within
    1e-4 of FP64, and the same bits on a second call."""
    torch = hip
    rng = np.random.default_rng(B * 1000 + H * 10 + Ci + Co)
    x = S.bf16(rng.standard_normal((B, H, W, Ci)))
    gz = S.bf16(rng.standard_normal((B, H, W, Co)))
    w = S.bf16(rng.standard_normal((Co, 3, 3, Ci)) * 0.05)
    _, dw, db = S.conv3x3_backward(x, w, gz)
    st = torch.cuda.current_stream().cuda_stream
    xd, gzd = to16(torch, x), to16(torch, gz)
    need = N.lib().bz_nn_wgrad_work(B, H, W, Ci, Co, 3)
    work = torch.empty(need, dtype=torch.float32).cuda()
    runs = []
    for _ in range(2):
        gw = torch.zeros((Co, 3, 3, Ci), dtype=torch.float32).cuda()
        gb = torch.full((Co,), 0.5, dtype=torch.float32).cuda()          # the sums add to what is there
        N.call("bz_nn_conv3x3_wgrad", xd.data_ptr(), gzd.data_ptr(), gw.data_ptr(), gb.data_ptr(), B, H, W, Ci,
               Co, 3, work.data_ptr(), need, st)
        runs.append((gw.cpu().numpy(), gb.cpu().numpy() - 0.5))
    assert rel_l2(runs[0][0], dw) < 1e-4, rel_l2(runs[0][0], dw)
    assert rel_l2(runs[0][1], db) < 1e-5, rel_l2(runs[0][1], db)
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("B,H,W,Ci,Co", [(2, 5, 128, 32, 32), (2, 3, 64, 64, 64), (2, 8, 16, 64, 128),
                                         (1, 4, 32, 32, 32)])
def test_fused_relu_derivatives_equal_the_separate_passes(hip, B, H, W, Ci, Co):
    """Backward-data with the ReLU mask of its forward input fused (row march, tiled GEMM and halo kernels), and
    the pooling backward with the mask of the level below fused: the same bits as the separate passes."""
    torch = hip
    rng = np.random.default_rng(B + H + W + Ci + Co)
    st = torch.cuda.current_stream().cuda_stream
    w = S.bf16(rng.standard_normal((Co, 3, 3, Ci)) * 0.05)
    w32 = torch.tensor(w, dtype=torch.float32).cuda()
    w16 = torch.empty((Co, 3, 3, Ci), dtype=torch.int16).cuda()
    wt16 = torch.empty((Ci, 3, 3, Co), dtype=torch.int16).cuda()
    N.call("bz_nn_weights", w32.data_ptr(), w16.data_ptr(), wt16.data_ptr(), Co, Ci, st)
    gzd = to16(torch, S.bf16(rng.standard_normal((B, H, W, Co))))
    xd = to16(torch, S.bf16(np.maximum(rng.standard_normal((B, H, W, Ci)), 0.0)))     # half of it zero
    a = torch.empty((B, H, W, Ci), dtype=torch.int16).cuda()
    b = torch.empty_like(a)
    N.call("bz_nn_conv3x3_dgrad", gzd.data_ptr(), wt16.data_ptr(), a.data_ptr(), B, H, W, Ci, Co, st)
    N.call("bz_nn_relu_mask", a.data_ptr(), xd.data_ptr(), a.numel(), st)
    N.call("bz_nn_conv3x3_dgrad_relu", gzd.data_ptr(), wt16.data_ptr(), xd.data_ptr(), b.data_ptr(), B, H, W, Ci, Co,
           st)
    assert torch.equal(a, b) and int((a == 0).sum()) > a.numel() // 4
    gy = to16(torch, S.bf16(rng.standard_normal((B, H // 2 or 1, W // 2, Ci)))) if H % 2 == 0 else None
    if gy is not None:
        p = torch.empty((B, H, W, Ci), dtype=torch.int16).cuda()
        q = torch.empty_like(p)
        N.call("bz_nn_avgpool2_bwd", gy.data_ptr(), p.data_ptr(), B, H, W, Ci, st)
        N.call("bz_nn_relu_mask", p.data_ptr(), xd.data_ptr(), p.numel(), st)
        N.call("bz_nn_avgpool2_bwd_relu", gy.data_ptr(), xd.data_ptr(), q.data_ptr(), B, H, W, Ci, st)
        assert torch.equal(p, q)


def _tree_f32(rows, out):
    """out + the rows in the reduction's fixed order, in FP32: groups of 16 consecutive rows summed in row order,
    level after level while more than 16 remain, then the last rows added to out in order."""
    cur = [np.asarray(r, np.float32) for r in rows]
    while len(cur) > 16:
        nxt = []
        for g in range(0, len(cur), 16):
            s = cur[g].copy()
            for r in cur[g + 1:g + 16]:
                s = (s + r).astype(np.float32)
            nxt.append(s)
        cur = nxt
    s = np.asarray(out, np.float32).copy()
    for r in cur:
        s = (s + r).astype(np.float32)
    return s


def _rows_for(need, per_row, extra_per_row):
    """Partial rows z of a workspace of `need` floats laid out as z rows of per_row floats plus ceil(z/16) scratch
    rows of extra_per_row (when z > 16)."""
    for z in range(1, 1 << 16):
        if z * per_row + (-(-z // 16) if z > 16 else 0) * extra_per_row == need:
            return z
    raise AssertionError("no row count fits the workspace size")


@pytest.mark.parametrize("B,H,W,Ci,Co,Z", [(2, 8, 16, 64, 128, 1), (16, 32, 32, 64, 128, 32),
                                           (128, 32, 32, 64, 128, 205), (128, 64, 64, 32, 64, 432),
                                           (128, 128, 128, 32, 32, 863)])
def test_weight_gradient_sums_follow_the_tree_order(hip, B, H, W, Ci, Co, Z):
    """The ordered reduction of the weight-gradient partials (one launch for all levels) gives the bits of the
    fixed tree over the Z partial rows the kernel left in the workspace, recomputed in FP32: no level, one level
    (32 and 205 rows), two levels (432 rows; 863 rows, where each element takes 32 lanes)."""
    torch = hip
    rng = np.random.default_rng(B + H + W + Ci + Co)
    st = torch.cuda.current_stream().cuda_stream
    xd = to16(torch, S.bf16(rng.standard_normal((B, H, W, Ci))))
    gzd = to16(torch, S.bf16(rng.standard_normal((B, H, W, Co))))
    need = N.lib().bz_nn_wgrad_work(B, H, W, Ci, Co, 3)
    nw = Co * 9 * Ci
    assert _rows_for(need, nw + Co, nw) == Z
    work = torch.empty(need, dtype=torch.float32).cuda()
    gw = torch.zeros(nw, dtype=torch.float32).cuda()
    gb = torch.full((Co,), 0.25, dtype=torch.float32).cuda()
    N.call("bz_nn_conv3x3_wgrad", xd.data_ptr(), gzd.data_ptr(), gw.data_ptr(), gb.data_ptr(), B, H, W, Ci, Co, 3,
           work.data_ptr(), need, st)
    wk = work.cpu().numpy()
    wpart = wk[:Z * nw].reshape(Z, nw)
    bpart = wk[Z * nw:Z * (nw + Co)].reshape(Z, Co)
    assert np.array_equal(gw.cpu().numpy(), _tree_f32(wpart, np.zeros(nw, np.float32)))
    assert np.array_equal(gb.cpu().numpy(), _tree_f32(bpart, np.full(Co, 0.25, np.float32)))


def test_first_layer_weight_gradient_sums_follow_the_tree_order(hip):
    """The same for the first layer's partial rows of 9 Co weight and Co bias columns (1024 rows, three
    levels)."""
    torch = hip
    rng = np.random.default_rng(3)
    st = torch.cuda.current_stream().cuda_stream
    B, H, W, Co = 64, 128, 128, 32
    xd = to16(torch, S.bf16(rng.random((B, H, W, 1))))
    gzd = to16(torch, S.bf16(rng.standard_normal((B, H, W, Co))))
    wd = torch.tensor(S.bf16(rng.standard_normal((Co, 9)) * 0.3), dtype=torch.float32).cuda()
    need = N.lib().bz_nn_first_work(B, H, W, Co)
    Z = _rows_for(need, 10 * Co, 10 * Co)
    work = torch.empty(need, dtype=torch.float32).cuda()
    gw = torch.zeros((Co, 9), dtype=torch.float32).cuda()
    gb = torch.zeros(Co, dtype=torch.float32).cuda()
    N.call("bz_nn_conv_first_bwd", xd.data_ptr(), gzd.data_ptr(), wd.data_ptr(), gw.data_ptr(), gb.data_ptr(), None,
           B, H, W, Co, work.data_ptr(), need, st)
    part = work.cpu().numpy()[:Z * 10 * Co].reshape(Z, 10 * Co)
    assert Z == 1024, Z
    assert np.array_equal(gw.cpu().numpy().ravel(), _tree_f32(part[:, :9 * Co], np.zeros(9 * Co, np.float32)))
    assert np.array_equal(gb.cpu().numpy(), _tree_f32(part[:, 9 * Co:], np.zeros(Co, np.float32)))


@pytest.mark.parametrize("B,H,W,Co", [(2, 32, 16, 32), (4, 16, 64, 64), (1, 32, 32, 16), (3, 8, 128, 32)])
def test_first_layer_weight_gradient_on_staged_rows_matches_oracle(hip, B, H, W, Co):
    """First-layer weight and bias gradients where a workgroup's chunk is whole rows of one image (the kernel that
    stages those rows of the input in LDS): 2, 4 and 8 channel groups, one image or several per launch, chunks of
    4 to 32 rows; through atomics and through the ordered partial sums, within 1e-5 of FP64. And the forward
    convolution, which stages rows the same way."""
    torch = hip
    rng = np.random.default_rng(B * 100 + H + W + Co)
    st = torch.cuda.current_stream().cuda_stream
    x = S.bf16(rng.random((B, H, W, 1)))
    w = S.bf16(rng.standard_normal((Co, 3, 3, 1)) * 0.3)
    gz = S.bf16(rng.standard_normal((B, H, W, Co)))
    _, dw, db = S.conv3x3_backward(x, w, gz)
    xd, gzd = to16(torch, x), to16(torch, gz)
    wd = torch.tensor(w.reshape(Co, 9), dtype=torch.float32).cuda()
    need = N.lib().bz_nn_first_work(B, H, W, Co)
    work = torch.empty(need, dtype=torch.float32).cuda()
    for wk, nwk in ((None, 0), (work.data_ptr(), need)):
        gw = torch.zeros((Co, 9), dtype=torch.float32).cuda()
        gb = torch.zeros(Co, dtype=torch.float32).cuda()
        N.call("bz_nn_conv_first_bwd", xd.data_ptr(), gzd.data_ptr(), wd.data_ptr(), gw.data_ptr(), gb.data_ptr(),
               None, B, H, W, Co, wk, nwk, st)
        assert rel_l2(gw.cpu().numpy(), dw.reshape(Co, 9)) < 1e-5, wk
        assert rel_l2(gb.cpu().numpy(), db) < 1e-5, wk
    # the forward on the same staged rows: within one bfloat16 step of FP64, and the bits of the per-element
    # kernel (an output address that is not 16-byte aligned takes that one)
    bias = rng.standard_normal(Co).astype(np.float32).astype(np.float64) * 0.1
    bd = torch.tensor(bias, dtype=torch.float32).cuda()
    y = torch.empty((B, H, W, Co), dtype=torch.int16).cuda()
    N.call("bz_nn_conv_first", xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), y.data_ptr(), B, H, W, Co, st)
    ref = np.maximum(S.conv3x3(x, w, bias), 0.0)
    assert np.all(np.abs(from16(y) - S.bf16(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    y2 = torch.empty(B * H * W * Co + 8, dtype=torch.int16).cuda()
    N.call("bz_nn_conv_first", xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), y2.data_ptr() + 2, B, H, W, Co, st)
    assert np.array_equal(y.cpu().numpy().ravel(), y2[1:1 + B * H * W * Co].cpu().numpy())
