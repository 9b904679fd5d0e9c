"""Total variation superiorization (NTVS_iteration, DS.cu:10583; iteratively_perturb_image :10476).

CPU part: the oracle pinned by what the mathematics fixes — closed-form total variations, the
subgradient against finite differences of the TV, the A^L table, the summation order by vectors
whose sum depends on it, and the accept/shrink loop against its definition.
GPU part: the library's TVS against the oracle's restatement in the library's summation order,
bit for bit (DROP made a no-op with LAMBDA = 0 so that only TVS acts), and with DROP interleaved.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from cases import small_head
from pct_reconstruction_amd import derive
from pct_reconstruction_amd.recon import default_params

SERIAL, TILED = 0, 1


def _params(C=12, R=10, S=3, **kw):
    # SLICES = (SSD_V_SIZE - 1) / SLICE_THICKNESS
    base = dict(recon_cyl_radius=10.0, columns=C, rows=R, slice_thickness=1.0, ssd_v_size=1.0 + S)
    base.update(kw)
    return derive(default_params(**base))


def _tv64(P, x):
    """TV in float64 (for finite differences): sum over rows < R-1, columns < C-1 of every slice."""
    a = x.reshape(P.slices, P.rows, P.columns).astype(np.float64)
    dy = a[:, 1:, :-1] - a[:, :-1, :-1]
    dx = a[:, :-1, 1:] - a[:, :-1, :-1]
    return float(np.sum(np.sqrt(dy * dy + dx * dx)))


# ---------------------------------------------------------------- summation order
def test_tile_sum_order_is_pinned_by_order_dependent_vectors(oracle):
    # 1e16 + 1 rounds back to 1e16 (ulp 2, tie to even): which ones meet first decides the result.
    big = 1e16
    # halving tree: threads 1 and 3 meet at stride 2 (1 + 1 = 2), then 1e16 + 2 is exact
    assert oracle.tile_sum(np.array([big, 1.0, 0.0, 1.0])) == big + 2
    # thread j adds voxels j, j+256, j+512, j+768: index 0, 256 and 768 meet inside thread 0
    t = np.zeros(1024)
    t[0], t[256], t[768] = big, 1.0, 1.0
    assert oracle.tile_sum(t) == big
    # tile sums: thread 1 of the final level adds tiles 1 and 257 (1 + 1) before meeting tile 0
    t = np.zeros(258 * 1024)
    t[0], t[1024], t[257 * 1024] = big, 1.0, 1.0
    assert oracle.tile_sum(t) == big + 2
    # and it is a sum
    rng = np.random.default_rng(3)
    r = rng.random(5000)
    assert oracle.tile_sum(r) == pytest.approx(math.fsum(r), rel=1e-14)
    assert oracle.tile_sum(np.arange(100000, dtype=np.float64)) == 99999 * 100000 / 2
    assert oracle.tile_sum(np.zeros(0)) == 0.0


# ---------------------------------------------------------------- total variation
@pytest.mark.parametrize("mode", [SERIAL, TILED])
def test_tv_of_a_ramp_and_of_spikes_is_closed_form(oracle, mode):
    P = _params()
    C, R, S = P.columns, P.rows, P.slices
    z, y, x = np.meshgrid(np.arange(S), np.arange(R), np.arange(C), indexing="ij")
    ramp = (3.0 * x + 4.0 * y + 7.0 * z).astype(np.float32).reshape(-1)  # every term sqrt(9 + 16) = 5
    assert oracle.total_variation(P, ramp, mode) == 5.0 * S * (R - 1) * (C - 1)

    def spike(c, r, s):
        a = np.zeros((S, R, C), np.float32)
        a[s, r, c] = 1.0
        return a.reshape(-1)

    # interior: its own term sqrt(2) and the terms of its left and upper neighbours (1 each)
    assert oracle.total_variation(P, spike(4, 5, 1), mode) == np.float32(2.0) + np.float32(math.sqrt(2.0))
    # last column / last row: the voxel's own term is skipped, only the forward difference into it counts
    assert oracle.total_variation(P, spike(C - 1, 5, 1), mode) == 1.0
    assert oracle.total_variation(P, spike(4, R - 1, 2), mode) == 1.0
    # corner (C-1, R-1): no term reaches it
    assert oracle.total_variation(P, spike(C - 1, R - 1, 0), mode) == 0.0


def test_serial_float_and_tiled_tv_agree_to_float_rounding(oracle):
    P = _params(C=40, R=36, S=4)
    x = np.random.default_rng(1).random(P.num_voxels).astype(np.float32)
    ref = _tv64(P, x)
    for mode in (SERIAL, TILED):
        assert oracle.total_variation(P, x, mode) == pytest.approx(ref, rel=1e-5)


# ---------------------------------------------------------------- perturbation
def test_perturbation_is_the_normalised_tv_gradient(oracle):
    P = _params(C=9, R=8, S=2)
    rng = np.random.default_rng(7)
    x = rng.random(P.num_voxels).astype(np.float32)
    G, v, norm = oracle.tvs_perturbation(P, x, TILED)
    # G = dTV/dx (TV is differentiable here: no zero gradient norm), by central differences
    x64 = x.astype(np.float64)
    fd = np.empty(P.num_voxels)
    h = 1e-6
    for i in range(P.num_voxels):
        xp, xm = x64.copy(), x64.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (_tv64(P, xp) - _tv64(P, xm)) / (2 * h)
    np.testing.assert_allclose(G, fd, rtol=1e-4, atol=1e-4)
    assert norm == pytest.approx(math.sqrt(np.sum(G.astype(np.float64) ** 2)), rel=1e-6)
    np.testing.assert_array_equal(v, (G / np.float32(norm)).astype(np.float32))
    assert np.sum(v.astype(np.float64) ** 2) == pytest.approx(1.0, rel=1e-6)
    # a constant image has G = 0 and v = 0 (norm 0 branch, DS.cu:10317-10320)
    G0, v0, n0 = oracle.tvs_perturbation(P, np.full(P.num_voxels, 2.5, np.float32), TILED)
    assert n0 == 0.0 and not G0.any() and not v0.any()


# ---------------------------------------------------------------- A^L table and the draw of L
def test_eta_table_and_draw(oracle):
    P = _params()
    assert oracle.tvs_eta_len(P) == 132  # (int)(log(1e-40) / log(0.5)) = (int)132.88
    for L in range(132):
        assert oracle.tvs_eta(P, L) == 2.0 ** -L
    assert oracle.tvs_eta(P, 132) == 0.0 and oracle.tvs_eta(P, 250) == 0.0
    Q = _params(tvs_l0=3, tvs_a=0.75)
    assert oracle.tvs_eta(Q, 2) == 0.0  # below L_0: no table entry
    assert oracle.tvs_eta(Q, 3) == np.float32(0.75 ** 3)  # entry L - L_0 holds A^L
    assert oracle.tvs_eta_len(Q) == int(math.log(1e-40) / math.log(0.75))
    draws = [oracle.tvs_draw(5, k, 10, 19) for k in range(1, 4001)]
    assert min(draws) == 10 and max(draws) == 19
    counts = np.bincount(np.array(draws) - 10, minlength=10)
    assert np.all(np.abs(counts - 400) < 80)
    assert oracle.tvs_draw(5, 7, 10, 19) == draws[6]
    assert oracle.tvs_draw(5, 7, 12, 12) == 12 and oracle.tvs_draw(5, 7, 12, 4) == 12


# ---------------------------------------------------------------- the accept / shrink loop
def _hull(P, seed=2):
    return (np.random.default_rng(seed).random(P.num_voxels) < 0.7).astype(np.uint8)


@pytest.mark.parametrize("mode", [SERIAL, TILED])
def test_conditioned_step_accepts_the_first_beta_that_does_not_raise_tv(oracle, mode):
    P = _params(C=16, R=14, S=3, tvs_old=1)  # one repetition, no draw
    hull = _hull(P)
    rng = np.random.default_rng(11)
    x = (rng.random(P.num_voxels) + np.linspace(0, 3, P.num_voxels)).astype(np.float32)
    y, L, tvb, tva = oracle.tvs_iteration(P, hull, x, 1, 0, mode)
    _, v, _ = oracle.tvs_perturbation(P, x, mode)
    tv_x = oracle.total_variation(P, x, mode)
    assert tvb == tv_x

    def cand(Lc):
        beta = np.float32(oracle.tvs_eta(P, Lc))
        return np.where(hull != 0, x - beta * v, x).astype(np.float32)

    Lb = L - 1  # the exponent of the accepted BETA (L was incremented once more after the last try)
    np.testing.assert_array_equal(y, cand(Lb))
    assert tva == oracle.total_variation(P, y, mode) <= tv_x
    for Lr in range(0, Lb):  # every larger BETA raised the TV
        assert oracle.total_variation(P, cand(Lr), mode) > tv_x
    assert np.array_equal(y[hull == 0], x[hull == 0])


def test_flat_image_stops_at_the_first_try(oracle):
    # v = 0 on a constant image, so every candidate has TV_y == TV_x: "TV_y <= TV_x" accepts at once,
    # one L increment per repetition (a strict "<" would run L up to 301)
    P = _params(C=16, R=14, S=3, tvs_seed=4)
    x = np.full(P.num_voxels, 1.25, np.float32)
    y, L, tvb, tva = oracle.tvs_iteration(P, _hull(P), x, 1, 0, TILED)
    assert L == P.tvs_repetitions and tvb == tva == 0.0
    np.testing.assert_array_equal(y, x)


def test_repetitions_draw_and_unconditioned_variant(oracle):
    P = _params(C=16, R=14, S=3, tvs_seed=9)
    hull = _hull(P, 4)
    x = np.random.default_rng(5).random(P.num_voxels).astype(np.float32)
    # k = 3 draws L in [L_0 + 2, max(L, L_0 + 2)]; from L = 0 that is 2
    y, L, tvb, tva = oracle.tvs_iteration(P, hull, x, 3, 0, TILED)
    assert L >= 2 + P.tvs_repetitions and tva <= tvb
    # five repetitions = five single-repetition steps at the same L sequence
    Q = _params(C=16, R=14, S=3, tvs_seed=9, tvs_repetitions=1, tvs_old=1)
    z, Lz = x, 2
    for _ in range(5):
        z, Lz, _, _ = oracle.tvs_iteration(Q, hull, z, 3, Lz, TILED)
    np.testing.assert_array_equal(y, z)
    assert Lz == L
    # unconditioned (iteratively_perturb_image_unconditional :10562): x -= A^L v, L += 1, per repetition
    U = _params(C=16, R=14, S=3, tvs_conditioned=0, tvs_old=1, tvs_repetitions=1)
    y1, L1, _, _ = oracle.tvs_iteration(U, hull, x, 1, 4, TILED)
    _, v, _ = oracle.tvs_perturbation(U, x, TILED)
    np.testing.assert_array_equal(y1, np.where(hull != 0, x - np.float32(2.0 ** -4) * v, x).astype(np.float32))
    assert L1 == 5


def test_serial_float_and_tiled_loops_take_the_same_decisions_on_a_small_image(oracle):
    P = _params(C=20, R=18, S=2, tvs_seed=1)
    hull = _hull(P, 6)
    x = np.random.default_rng(8).random(P.num_voxels).astype(np.float32)
    a, La, _, _ = oracle.tvs_iteration(P, hull, x, 1, 0, SERIAL)
    b, Lb, _, _ = oracle.tvs_iteration(P, hull, x, 1, 0, TILED)
    assert La == Lb
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-6)


# ---------------------------------------------------------------- GPU: the library against the oracle
def _tvs_only_case(**kw):
    # LAMBDA = 0 makes every DROP update exactly 0, so x changes only through TVS.
    base = dict(n=20000, seed=3, lambda_=0.0, tvs_on=1, iterations=3, tvs_seed=17)
    base.update(kw)
    return small_head(**base)


def _oracle_tvs_sequence(oracle, P, hull, x, iterations):
    L, log = 0, []
    for k in range(1, iterations + 1):
        x, L, tb, ta = oracle.tvs_iteration(P, hull, x, k, L, TILED)
        log.append((tb, ta, L))
    return x, log


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [dict(), dict(tvs_first=0), dict(tvs_conditioned=0), dict(tvs_old=1),
                                     dict(tvs_repetitions=2, tvs_a=0.7, tvs_l0=1)])
def test_gpu_tvs_is_bit_exact_with_the_oracle(variant):
    from pct_reconstruction_amd import Reconstruction
    import pct_oracle as oracle
    P, data = _tvs_only_case(**variant)
    print(f"[tvs {variant}] histories {data['wepl'].shape[0]}", flush=True)
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.import_and_process_data()
        r.hull_detection_finish()
        r.statistical_calculations_and_cuts()
        r.generate_preprocessing_images()
        r.reconstruction_cuts()
        r.build_mlp_lists()
        hull = r.image("hull").reshape(-1).astype(np.uint8)
        rng = np.random.default_rng(21)
        x0 = (r.image("x0").reshape(-1) + 0.05 * rng.standard_normal(P.num_voxels)).astype(np.float32)
        r.set_image(x0)
        assert r.total_variation() == oracle.tile_sum(_tv_terms(P, x0))
        r.DROP(P.iterations)
        x = r.image("x").reshape(-1)
        log = r.tvs_log()
    ref, ref_log = _oracle_tvs_sequence(oracle, P, hull, x0, P.iterations)
    print(f"[tvs {variant}] L {[e[2] for e in log]} oracle {[e[2] for e in ref_log]}", flush=True)
    assert [e[2] for e in log] == [e[2] for e in ref_log]
    assert [(e[0], e[1]) for e in log] == [(e[0], e[1]) for e in ref_log]
    assert np.array_equal(x, ref)
    assert not np.array_equal(x, x0)


def _tv_terms(P, x):
    a = x.reshape(P.slices, P.rows, P.columns)
    t = np.zeros_like(a)
    dy = a[:, 1:, :-1] - a[:, :-1, :-1]
    dx = a[:, :-1, 1:] - a[:, :-1, :-1]
    t[:, :-1, :-1] = np.sqrt(dy * dy + dx * dx)
    return t.reshape(-1).astype(np.float64)


@pytest.mark.gpu
def test_gpu_pipeline_with_tvs_between_drop_iterations_matches_the_oracle():
    from pct_reconstruction_amd import Reconstruction
    import pct_oracle as oracle
    P, data = small_head(n=40000, seed=9, tvs_on=1, iterations=3, tvs_seed=2)
    ref = oracle.run_pipeline(P, data, iterations=P.iterations, accum=P.drop_accum)
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        x = r.run().reshape(-1)
        log = r.tvs_log()
        m = r.metrics()
    rel = np.linalg.norm(x - ref["x"]) / np.linalg.norm(ref["x"])
    print(f"[tvs pipeline] rel L2 {rel:.3e}; L {[e[2] for e in log]} oracle {[e[2] for e in ref['tvs_log']]}; "
          f"tvs {m['tvs']}", flush=True)
    assert [e[2] for e in log] == [e[2] for e in ref["tvs_log"]]
    for (a, b, _), (c, d, _) in zip(log, ref["tvs_log"]):
        assert a == pytest.approx(c, rel=1e-5) and b == pytest.approx(d, rel=1e-5)
    assert rel <= 1e-5
    assert m["tvs"]["iterations"] == P.iterations
