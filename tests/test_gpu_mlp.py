"""Learned Newton guess on the GPU: data, FP32 forward/backward against the
FP64 numpy oracle, BF16 matrix-core inference against the FP32 forward,
training, and what the trained network saves the exact solver."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from tolerances import tolerance_ratio

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402

import finitevolumesolver_amd as fvs  # noqa: E402
from finitevolumesolver_amd.newton_guess import LearnedNewtonGuess, exact_middle_state  # noqa: E402

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def data(cuda):
    net = LearnedNewtonGuess(seed=21)
    counts = net.generate(8192, seed=5)
    return net, counts, net.dataset()


def _median_abs_error(net, hd):
    ok = hd["flag"] == 0
    y = net.forward(hd["features"])
    return float((y - hd["target"])[ok].abs().median().item())


@pytest.fixture(scope="module")
def trained(cuda):
    import time

    import torch
    net = LearnedNewtonGuess(seed=1)
    counts = net.generate(1 << 18, seed=2)
    held = LearnedNewtonGuess(seed=0)
    held.generate(16384, seed=99)
    hd = held.dataset()
    e0 = _median_abs_error(net, hd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.train(epochs=32, batch=1024, lr=5e-3, huber=0.25, seed=3)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    e1 = _median_abs_error(net, hd)
    _note(f"training: counts {counts}, {secs:.2f} s, held-out median |r error| "
          f"{e0:.4f} -> {e1:.4f}, epoch losses {np.round(hist, 4).tolist()}")
    return net, hd, e0, e1, hist


def test_dataset_matches_oracle(data):
    net, counts, ds = data
    prim = ds["prim"].cpu().numpy()
    n = prim.shape[1]
    assert counts["usable"] + counts["vacuum"] + counts["not_converged"] == n
    assert counts["not_converged"] == 0
    # the generator (SplitMix64 uniforms, log-uniform rho and p)
    ref = M.sample_prim(5, range(256))
    np.testing.assert_allclose(prim[:, :256], ref, rtol=1e-14, atol=0)
    np.testing.assert_allclose(ds["features"].cpu().numpy(), M.features(prim), rtol=0, atol=2e-6)
    flag = ds["flag"].cpu().numpy()
    vac = M.vacuum(prim)
    assert np.array_equal(flag == 1, vac)
    assert counts["vacuum"] == int(vac.sum())
    ok = flag == 0
    left, right = M.complete_pairs(prim[:, ok])
    pM, _, it, conv = O.exact_middle_state(left, right, 2, 0)
    assert np.all(conv == 1)
    r = np.log(pM) - 0.5 * (np.log(prim[2, ok]) + np.log(prim[5, ok]))
    t = ds["target"].cpu().numpy()[ok]
    np.testing.assert_allclose(t, r, rtol=0, atol=1e-6 * (1 + np.abs(r)).max())
    gpu_it = ds["iterations"].cpu().numpy()[ok]
    same = np.mean(gpu_it == it)
    _note(f"dataset: {ok.sum()} usable, iterations identical for {same:.4%}, "
          f"mean {gpu_it.mean():.2f}")
    assert same >= 0.99


def test_fp32_forward_and_gradients_match_oracle(data):
    import torch
    net, _, ds = data
    p = net.get_params().astype(np.float64)
    x = ds["features"][:, :2048].contiguous()
    t = ds["target"][:2048].contiguous()
    flag = ds["flag"][:2048].contiguous()
    y = net.forward(x).cpu().numpy()
    xr = x.cpu().numpy().astype(np.float64)
    yr, _, _ = M.forward(p, xr)
    err = np.max(np.abs(y - yr) / (1 + np.abs(yr)))
    _note(f"fp32 forward vs fp64 oracle: max err {err:.2e}")
    assert err <= 2e-5
    for huber in (0.0, 0.25):
        g, loss = net.gradients(x, t, flag, huber=huber)
        lr_, gr = M.loss_and_grad(p, xr, t.cpu().numpy(), flag.cpu().numpy(), huber=huber)
        assert loss == pytest.approx(lr_, rel=1e-4)
        gerr = np.max(np.abs(g - gr)) / np.max(np.abs(gr))
        _note(f"gradient (huber {huber}) vs fp64 oracle: max |dg| / max |g| = {gerr:.2e}")
        assert gerr <= 1e-4
    # every parameter block carries gradient, padding none
    for lo, hi in [(M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_W3, M.OFF_B3)]:
        assert np.any(g[lo:hi] != 0)
    assert np.all(g[M.OFF_W1:M.OFF_B1].reshape(64, 8)[:, 6:] == 0)
    torch.cuda.synchronize()


def _pairs_on_device(prim, torch, dev):
    left, right = M.complete_pairs(prim)
    return torch.from_numpy(left).to(dev), torch.from_numpy(right).to(dev)


def test_matrix_core_inference_matches_fp32(data, cuda):
    """BF16 MFMA inference against the FP32 forward of the same parameters:
    a wrong operand layout moves outputs by O(1); bf16 rounding by ~1e-2."""
    import torch
    net, _, ds = data
    rng = np.random.default_rng(4)
    p = net.get_params()
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.2, 64)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.2, 64)
    p[M.OFF_B3] = 0.3
    net.set_params(p)
    prim = ds["prim"].cpu().numpy()[:, :4096]
    for dirn in (0, 1):
        left, right = M.complete_pairs(prim)
        if dirn == 1:  # same pairs with the normal momentum in the y slot
            left, right = left[[0, 2, 1, 3, 4, 5]], right[[0, 2, 1, 3, 4, 5]]
        tl, tr = torch.from_numpy(left).to(cuda), torch.from_numpy(right).to(cuda)
        r, guess = net.infer_pairs(2, dirn, tl, tr)
        y = net.forward(ds["features"][:, :4096].contiguous())
        r, y, guess = r.cpu().numpy(), y.cpu().numpy(), guess.cpu().numpy()
        d = np.abs(r - y)
        _note(f"mfma vs fp32 (dir {dirn}): max {d.max():.3e} mean {d.mean():.3e}, "
              f"|y| mean {np.abs(y).mean():.3f}")
        assert d.max() <= 0.08 and d.mean() <= 0.02
        g_ref = np.sqrt(prim[2] * prim[5]) * np.exp(r.astype(np.float64))
        # weak pairs keep the mean-pressure guess
        a_l, a_r = left[5], right[5]
        sp = prim[2] + prim[5]
        weak = ((np.abs(prim[2] - prim[5]) < 0.1 * sp)
                & (np.abs(prim[4] - prim[1]) * (prim[0] + prim[3]) * (a_l + a_r) * 0.25 < 0.1 * sp))
        g_ref = np.where(weak, 0.5 * sp, g_ref)
        # exp of the network output comes from the hardware exp2 estimate
        np.testing.assert_allclose(guess, g_ref, rtol=1e-6)
        np.testing.assert_array_equal(guess[weak], 0.5 * sp[weak])


def test_face_guess_layout(data, cuda):
    """Patch face guesses equal pair inference on the explicitly built
    (clamped) neighbour pairs of every face, in the face-box layout."""
    import torch
    from fields import rough_field
    net, _, _ = data
    for shape in [(1, 9, 13), (5, 6, 7)]:
        D = 2 if shape[0] == 1 else 3
        cons = rough_field(shape, seed=81)
        q = O.complete_from_cons(cons.reshape(D + 2, -1), D).reshape((D + 4,) + shape)
        tc = torch.from_numpy(cons).to(cuda)
        for d in range(D):
            ax = 3 - d  # numpy axis of direction d in (comp, z, y, x)
            first = np.take(q, [0], axis=ax)
            last = np.take(q, [q.shape[ax] - 1], axis=ax)
            left = np.concatenate([first, q], axis=ax)
            right = np.concatenate([q, last], axis=ax)
            _, g_ref = net.infer_pairs(D, d, torch.from_numpy(left.reshape(D + 4, -1).copy()).to(cuda),
                                       torch.from_numpy(right.reshape(D + 4, -1).copy()).to(cuda))
            g = net.face_guess(D, 1.4, tc, d)
            fshape = list(shape)
            fshape[2 - d] += 1
            assert tuple(g.shape) == tuple(fshape)
            np.testing.assert_allclose(g.cpu().numpy().ravel(), g_ref.cpu().numpy(), rtol=1e-14)


def test_training_reduces_error(trained):
    _, _, e0, e1, hist = trained
    assert np.all(np.isfinite(hist))
    assert e1 < 0.05 and e1 < 0.1 * e0
    assert hist[-1] < hist[0]


def test_trained_guess_saves_newton_iterations(trained, cuda):
    """On held-out pairs the learned guess converges to the same star
    pressure in fewer Newton iterations than the reference's mean guess."""
    import torch
    net, hd, _, _, _ = trained
    ok = (hd["flag"] == 0).cpu().numpy()
    prim = hd["prim"].cpu().numpy()[:, ok]
    left, right = _pairs_on_device(prim, torch, cuda)
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=2), base="Godunov", reconstruction="None")
    p0, _, it0, c0 = exact_middle_state(fm, 0, left, right)
    _, guess = net.infer_pairs(2, 0, left, right)
    p1, _, it1, c1 = exact_middle_state(fm, 0, left, right, guess)
    it0, it1 = it0.cpu().numpy(), it1.cpu().numpy()
    p0, p1 = p0.cpu().numpy(), p1.cpu().numpy()
    assert np.all(c0.cpu().numpy() == 1)
    conv1 = c1.cpu().numpy() == 1
    _note(f"held-out {ok.sum()} pairs: mean iterations {it0.mean():.3f} (mean guess) -> "
          f"{it1.mean():.3f} (learned); max {it0.max()} -> {it1.max()}; "
          f"non-converged {np.sum(~conv1)}")
    assert np.mean(conv1) >= 0.999
    assert it1[conv1].mean() < 0.6 * it0.mean()
    # both stop at |f(p)| <= 1e-7, which fixes the star pressure only to
    # about f / f'(p): the two roots agree to that accuracy
    rel = np.abs(p1 - p0)[conv1] / p0[conv1]
    _note(f"star pressure agreement: median {np.median(rel):.1e}, max {rel.max():.1e}")
    assert np.median(rel) <= 1e-8 and rel.max() <= 1e-3
    # the oracle labels agree with the mean-guess GPU roots
    pr, _, _, _ = O.exact_middle_state(*M.complete_pairs(prim), 2, 0)
    np.testing.assert_allclose(p0, pr, rtol=1e-12)


def _blocky_field(n, seed, block=3):
    """3D piecewise-constant blocks with log-uniform density and pressure in
    [0.1, 10] and velocities in [-1, 1]: nearly every block face is a strong
    Riemann problem."""
    from finitevolumesolver_amd.initial_data import SplitMix64, cons_from_prim
    rng = SplitMix64(seed)
    nb = (n + block - 1) // block
    def field(lo, hi, log):
        v = rng.uniform(nb ** 3).reshape(nb, nb, nb)
        v = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * v) if log else lo + (hi - lo) * v
        return np.repeat(np.repeat(np.repeat(v, block, 0), block, 1), block, 2)[:n, :n, :n].copy()
    rho, p = field(0.1, 10.0, True), field(0.1, 10.0, True)
    vel = [field(-1.0, 1.0, False) for _ in range(3)]
    return cons_from_prim(rho, vel, p, 1.4)


def _level_runs(net, cons, n, dx, steps, cuda):
    import torch
    runs = {}
    for learned in (False, True):
        fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), base="Godunov", reconstruction="None")
        if learned:
            net.attach(fm)
        lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx, splitting="Godunov", cfl=0.8)
        lvl.set_state(torch.from_numpy(cons).to(cuda))
        fm.newton_stats(reset=True)
        done, fails, dts = lvl.advance(steps, batch=steps, return_dts=True)
        assert done == steps
        runs[learned] = (lvl.get_state().cpu().numpy(), fm.newton_stats(), np.array(dts))
    ref, ref_dts, _ = O.level_run(fvs.make_desc(fvs.PerfectGas(dim=3), base="Godunov",
                                                reconstruction="None"), cons, dx, steps)
    return runs, ref


def _l1_density(got, ref):
    return float(np.sum(np.abs(got[0] - ref[0])) / np.sum(np.abs(ref[0])))


def test_level_with_learned_guess_smooth(trained, cuda):
    """3D first-order exact Godunov on the smooth Fourier-mode field, 24^3,
    6 steps: nearly every face is a weak pair, so the learned form keeps the
    mean guess there; L1 density error <= 1e-6 against the FP64 oracle."""
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    n = 24
    cons, dx = fourier_modes_3d(n, seed0=12)
    runs, ref = _level_runs(trained[0], cons, n, dx, 6, cuda)
    (base, st0, _), (got, st1, _) = runs[False], runs[True]
    l1 = _l1_density(got, ref)
    linf = float(np.max(np.abs(got[0] - ref[0]) / np.abs(ref[0])))
    _note(f"smooth level: L1 density {l1:.2e}, Linf density {linf:.2e} (learned); "
          f"FP64 tolerance ratio {tolerance_ratio(base, ref):.2f} (mean guess); Newton "
          f"iterations {st0['iterations']} -> {st1['iterations']} over {st1['faces']} faces")
    assert l1 <= 1e-6
    assert tolerance_ratio(base, ref) <= 1.0
    assert st1["not_converged"] == 0
    assert st1["iterations"] <= 1.02 * st0["iterations"]


def test_level_with_learned_guess_strong_waves(trained, cuda):
    """3D first-order exact Godunov on blocky data with strong jumps, 24^3,
    4 steps: the learned guess saves Newton iterations (only waves with at
    least 48 non-weak faces run the network, so about a tenth here); L1
    density error <= 1e-6 against the FP64 oracle."""
    n = 24
    cons = _blocky_field(n, seed=5)
    dx = (1.0 / n, 1.0 / n, 1.0 / n)
    runs, ref = _level_runs(trained[0], cons, n, dx, 4, cuda)
    (base, st0, _), (got, st1, _) = runs[False], runs[True]
    l1 = _l1_density(got, ref)
    linf = float(np.max(np.abs(got[0] - ref[0]) / np.abs(ref[0])))
    _note(f"strong-wave level: L1 density {l1:.2e}, Linf density {linf:.2e} "
          f"(learned); Newton iterations {st0['iterations']} -> {st1['iterations']} over "
          f"{st1['faces']} faces, fallbacks {st1['fallbacks']}")
    assert l1 <= 1e-6
    assert st1["not_converged"] == 0
    assert st1["iterations"] < 0.93 * st0["iterations"]


@pytest.mark.parametrize("field", ["smooth", "strong"])
def test_level_mixed_precision_learned_form(trained, cuda, field):
    """The mixed-precision learned form of first-order exact Godunov
    (precision "fp32": FP64 Newton seeded by the network with exp/log
    powers, float sample and flux, FP64 update): L1 density error <= 1e-6
    against the FP64 oracle on smooth and strong-wave data, every Newton
    iteration converged."""
    import torch
    from finitevolumesolver_amd.initial_data import fourier_modes_3d
    n = 24
    if field == "smooth":
        cons, dx = fourier_modes_3d(n, seed0=12)
        steps = 6
    else:
        cons = _blocky_field(n, seed=5)
        dx = (1.0 / n, 1.0 / n, 1.0 / n)
        steps = 4
    fm = fvs.FluxMethod(fvs.PerfectGas(dim=3), base="Godunov", reconstruction="None",
                        precision="fp32")
    trained[0].attach(fm)
    lvl = fvs.SinglePatchLevel(fm, (n, n, n), dx, splitting="Godunov", cfl=0.8)
    lvl.set_state(torch.from_numpy(cons).to(cuda))
    fm.newton_stats(reset=True)
    done, fails = lvl.advance(steps, batch=steps)
    assert done == steps
    got = lvl.get_state().cpu().numpy()
    st = fm.newton_stats()
    ref, _, _ = O.level_run(fvs.make_desc(fvs.PerfectGas(dim=3), base="Godunov",
                                          reconstruction="None"), cons, dx, steps)
    l1 = _l1_density(got, ref)
    linf = float(np.max(np.abs(got[0] - ref[0]) / np.abs(ref[0])))
    _note(f"mixed learned ({field}): L1 density {l1:.2e}, Linf density {linf:.2e}; "
          f"{st['iterations']} Newton iterations over {st['faces']} faces, "
          f"fallbacks {st['fallbacks']}")
    assert l1 <= 1e-6
    assert st["not_converged"] == 0
