"""Learned warm-start network: the FP64 CPU oracle pinned by finite differences, and the gfx950
BF16/MFMA forward, backward and training checked against it."""
from __future__ import annotations

import numpy as np
import pytest


def _problem(shape, seed):
    D, H, W = shape
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    r2 = ((y - H / 2 + 0.5) / (H / 2)) ** 2 + ((x - W / 2 + 0.5) / (W / 2)) ** 2
    hull = (r2 < 0.8).astype(np.uint8)
    fbp = (hull * (1.0 + 0.3 * np.sin(0.7 * x + 0.3 * z) + 0.1 * rng.standard_normal(shape))).astype(np.float32)
    target = (hull * (1.0 + 0.25 * np.sin(0.7 * x + 0.3 * z))).astype(np.float32)
    return fbp, hull, target


def _params(oracle, seed, head_scale=0.2):
    """He-normal weights, small random biases and head (so every layer carries gradient)."""
    n = oracle.net_num_params()
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    cin = [8, 16, 32, 64, 192, 96, 48]
    cout = [16, 32, 64, 128, 64, 32, 16]
    off = 0
    for ci, co in zip(cin, cout):
        nw = co * 27 * ci
        p[off:off + nw] = rng.standard_normal(nw) * np.sqrt(2.0 / (27 * ci))
        off += nw
        p[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    p[off:off + 17] = head_scale * rng.standard_normal(17)
    assert off + 17 == n
    return p


def _layer_ranges():
    cin = [8, 16, 32, 64, 192, 96, 48]
    cout = [16, 32, 64, 128, 64, 32, 16]
    out, off = [], 0
    for ci, co in zip(cin, cout):
        out.append((off, off + co * 27 * ci))
        off += co * 27 * ci + co
    out.append((off, off + 17))
    return out


def test_net_oracle_gradient_matches_finite_differences(oracle):
    shape = (8, 8, 8)
    fbp, hull, target = _problem(shape, 1)
    p = _params(oracle, 2)
    loss, g, _ = oracle.net_loss_grad(shape, p, fbp, hull, target, bf16=False)
    assert loss > 0
    rng = np.random.default_rng(3)
    picks = []
    for a, b in _layer_ranges():
        seg = np.abs(g[a:b])
        picks.append(a + int(np.argmax(seg)))          # the largest gradient of each layer
        picks.append(a + int(rng.integers(0, b - a)))  # and a random one
    # small steps: a ReLU kink between p - eps and p + eps bends the finite difference (at 1e-3 a few
    # per cent on this volume), at 1e-5 the central difference agrees to ~1e-10
    for j in picks:
        eps = np.float32(1e-5) * max(np.float32(1.0), abs(p[j]))
        pp, pm = p.copy(), p.copy()
        pp[j] = p[j] + eps
        pm[j] = p[j] - eps
        dp, dm = float(pp[j]) - float(p[j]), float(p[j]) - float(pm[j])
        lp = oracle.net_loss_grad(shape, pp, fbp, hull, target, with_grad=False)[0]
        lm = oracle.net_loss_grad(shape, pm, fbp, hull, target, with_grad=False)[0]
        fd = (lp - lm) / (dp + dm)
        assert abs(fd - g[j]) <= 1e-6 * max(abs(g[j]), 1e-3 * np.abs(g).max()), (j, fd, g[j])


def test_net_oracle_zero_head_is_hybrid_fbp(oracle):
    shape = (8, 8, 8)
    fbp, hull, target = _problem(shape, 4)
    p = _params(oracle, 5, head_scale=0.0)
    loss, _, x0 = oracle.net_loss_grad(shape, p, fbp, hull, target, with_grad=False)
    np.testing.assert_array_equal(x0, (hull * fbp).astype(np.float64))
    m = hull.astype(bool)
    assert abs(loss - np.mean((fbp[m].astype(np.float64) - target[m]) ** 2)) < 1e-12


@pytest.mark.gpu
def test_net_gpu_forward_backward_match_oracle(oracle):
    from pct_reconstruction_amd.learned import WarmStartNet
    shape = (8, 16, 16)
    fbp, hull, target = _problem(shape, 6)
    p = _params(oracle, 7)
    ref_loss, ref_g, ref_x0 = oracle.net_loss_grad(shape, p, fbp, hull, target, bf16=True)
    f64_loss, f64_g, _ = oracle.net_loss_grad(shape, p, fbp, hull, target, bf16=False)
    with WarmStartNet(shape[2], shape[1], shape[0], seed=0) as net:
        assert net.num_params == p.size
        net.set_params(p)
        np.testing.assert_array_equal(net.get_params(), p)
        loss, g = net.loss_grad(fbp, hull, target)
        x0 = net.predict(fbp, hull)
    print(f"loss gpu {loss:.6e} oracle-bf16 {ref_loss:.6e} oracle-fp64 {f64_loss:.6e}", flush=True)
    assert abs(loss - ref_loss) <= 2e-3 * ref_loss
    assert np.linalg.norm(x0 - ref_x0) <= 2e-3 * np.linalg.norm(ref_x0)
    for k, (a, b) in enumerate(_layer_ranges()):
        rel = np.linalg.norm(g[a:b] - ref_g[a:b]) / np.linalg.norm(ref_g[a:b])
        rel64 = np.linalg.norm(g[a:b] - f64_g[a:b]) / np.linalg.norm(f64_g[a:b])
        print(f"layer {k}: rel L2 vs oracle-bf16 {rel:.2e}, vs fp64 {rel64:.2e}", flush=True)
        assert rel <= 2e-2, k
        assert rel64 <= 1e-1, k


@pytest.mark.gpu
def test_net_training_reduces_loss_and_weights_round_trip(oracle, tmp_path):
    from pct_reconstruction_amd.learned import WarmStartNet
    shape = (16, 32, 32)
    with WarmStartNet(shape[2], shape[1], shape[0], seed=1) as net:
        for s in range(4):
            net.add_sample(*_problem(shape, 10 + s))
        assert net.num_samples == 4
        fbp, hull, target = _problem(shape, 10)
        l0, _ = net.loss_grad(fbp, hull, target, with_grad=False)
        net.train(300, lr=1e-3, seed=0)
        l1, _ = net.loss_grad(fbp, hull, target, with_grad=False)
        print(f"loss before {l0:.4e} after {l1:.4e}", flush=True)
        assert l1 < 0.5 * l0
        path = tmp_path / "w.bin"
        net.save(path)
        x_a = net.predict(fbp, hull)
        with WarmStartNet(shape[2], shape[1], shape[0], seed=9) as net2:
            net2.load(path)
            np.testing.assert_array_equal(net2.predict(fbp, hull), x_a)


@pytest.mark.gpu
def test_committed_weights_reach_the_fbp_start_residual_in_two_iterations():
    """The shipped net (learned-128, input = the iterate after one DROP iteration) on a held-out
    phantom: the net's image lowers ||Ax - b|| below the FBP start's 20-iteration residual, so the
    learned start needs 1 + 1 iterations instead of 20."""
    from pathlib import Path

    from pct_reconstruction_amd import configs, synth
    from pct_reconstruction_amd.learned import WarmStartNet
    from pct_reconstruction_amd.recon import Reconstruction

    weights = Path(__file__).resolve().parent.parent / "pct_reconstruction_amd" / "weights" / \
        "warmstart_learned128_x1.bin"
    P, ph, bm, n = configs.learned_128(100019, 10_000_000)
    d = synth.generate(ph, bm, 100019, n)
    with Reconstruction(P) as r, WarmStartNet(P.columns, P.rows, P.slices, seed=0) as net:
        net.load(str(weights))
        r.load_histories(d["t"], d["v"], d["u"], d["wepl"], d["angle_deg"])
        r.preprocessing()
        r.reconstruction_cuts()
        r.build_mlp_lists()
        r.DROP(20)
        r_ref = r.residual()
        r.reset_iterate()
        r.DROP(1)
        j = net.apply_iterate(r)
        r.DROP(1)
        res = r.residual()
    print(f"[warm start] r_ref {r_ref:.2f}; x1 {j['residual_x']:.2f} -> net {j['residual_net']:.2f}; "
          f"after one more iteration {res:.2f}", flush=True)
    assert j["used_net"] and j["residual_net"] < j["residual_x"]
    assert res <= r_ref
