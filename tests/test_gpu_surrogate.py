"""GPU tests of the learned return-map surrogate (BF16 MFMA MLP) through the C ABI.

  forward            GPU predictions against oracle/mlp_oracle.py with the kernels' BF16 roundings
                     emulated: |dy| <= 2e-3 * max(1, max|y|) on the normalised outputs; a one-path
                     chain network against its closed form (BF16 weights), |dy| <= 1e-2
  gradient           GPU gradient against the emulated oracle: relative L2 error <= 1e-2, cosine
                     >= 0.9999, loss relative error <= 1e-4
  training           bitwise deterministic for a fixed seed; the loss falls by 10x or more
  pair logging       logged (trial, projected) pairs against the FP64 oracle projection
                     (oracle/mpm_oracle.cpp dp_project): |dH| <= 2e-6
  gate               a network whose outputs are NaN is rejected everywhere and the substep is
                     bitwise equal to the exact return map; a trained network is accepted for some
                     particles and the positions stay within 1e-4 of the exact run; the device's
                     accept flags equal the oracle gate's (oracle/mlp_oracle.gate) except within
                     1e-8 of a threshold, and its predictions the oracle network's within 1e-3 of
                     the forward bar
  gains              on held-out pairs the trained network is accepted for at least half, the
                     untrained network for at most a tenth as many, no network (H = eps) for none;
                     its mean error is at most half of either's
"""
import sys

import numpy as np
import pytest

from anisotropicelastoplasticity_amd import Solver, _lib, scenes
from oracle import bind as ob
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(n, seed=0):
    rng = np.random.default_rng(seed)
    inp = np.empty((n, 5), np.float32)
    e = np.sort(rng.normal(0.0, 0.01, size=(n, 3)), axis=1)[:, ::-1]
    inp[:, :3] = e
    inp[:, 3] = rng.uniform(25.0, 40.0, size=n)
    inp[:, 4] = rng.uniform(1.5, 3.5, size=n)
    return inp


def _solver():
    s = Solver()
    return s


def test_surrogate_forward_matches_oracle():
    n = 4099  # not a multiple of 16 or 64: ragged last wave and workgroup
    inp = _inputs(n, 1)
    with _solver() as s:
        assert s.surrogate_nparam() == mo.NPARAM
        s.surrogate_init(123)
        w = s.surrogate_get_weights()
        assert np.all(w[~mo.active_mask()] == 0.0)
        say(f"[surrogate forward] n={n}")
        H = s.surrogate_predict(inp)
    y_gpu = (H.astype(np.float64) - inp[:, :3]) / mo.CORR_SCALE
    y_ref, _ = mo.forward(w, inp, emulate_bf16=True)
    err = np.max(np.abs(y_gpu - y_ref))
    say(f"  max|dy| = {err:.3e}, max|y| = {np.max(np.abs(y_ref)):.3e}")
    # H = eps + 1e-3 y is formed in FP32: allow its rounding on top of the network difference
    assert err <= 2e-3 * max(1.0, np.max(np.abs(y_ref))) + 1e-3


def test_surrogate_forward_chain_closed_form():
    w = np.zeros(mo.NPARAM, np.float32)
    path = [(7, 3), (11, 7), (2, 11), (100, 2), (1, 100)]
    c = [0.9, -1.3, 0.7, 1.1, 2.0]
    bias = [0.1, -0.2, 0.05, 0.3, -0.4]
    for l, ((o, k), cl, bl) in enumerate(zip(path, c, bias)):
        w[mo.weight_offset(l) + o * mo.K[l] + k] = cl
        w[mo.bias_offset(l) + o] = bl
    inp = _inputs(64, 2)
    with _solver() as s:
        s.surrogate_set_weights(w)
        H = s.surrogate_predict(inp)
    y_gpu = (H.astype(np.float64) - inp[:, :3]) / mo.CORR_SCALE
    y_plain, _ = mo.forward(w, inp, emulate_bf16=False)
    y_emul, _ = mo.forward(w, inp, emulate_bf16=True)
    assert np.max(np.abs(y_gpu[:, [0, 2]])) <= 1e-3  # only output 1 is connected
    assert np.max(np.abs(y_gpu - y_emul)) <= 2e-3
    assert np.max(np.abs(y_gpu - y_plain)) <= 1e-2 * max(1.0, np.max(np.abs(y_plain)))


def test_surrogate_gradient_matches_oracle():
    n = 1000
    inp = _inputs(n, 3)
    rng = np.random.default_rng(4)
    H = inp[:, :3] + 1e-3 * rng.normal(0.0, 1.0, size=(n, 3)).astype(np.float32)
    with _solver() as s:
        s.surrogate_init(77)
        w = s.surrogate_get_weights()
        say(f"[surrogate gradient] n={n}")
        g_gpu, loss_gpu = s.surrogate_gradient(inp, H)
    loss_ref, g_ref = mo.loss_and_gradient(w, inp, H, emulate_bf16=True)
    relerr = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    cos = float(g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref)))
    say(f"  loss {loss_gpu:.6e} vs {loss_ref:.6e}; grad rel err {relerr:.3e}, cos {cos:.6f}")
    assert abs(loss_gpu - loss_ref) <= 1e-4 * loss_ref
    assert relerr <= 1e-2 and cos >= 0.9999
    assert np.all(g_gpu[~mo.active_mask()] == 0.0)
    # per-layer check so that no layer hides behind the others
    for l in range(5):
        sl = slice(mo.weight_offset(l), mo.bias_offset(l) + mo.N[l])
        r = np.linalg.norm(g_gpu[sl] - g_ref[sl]) / max(np.linalg.norm(g_ref[sl]), 1e-30)
        assert r <= 2e-2, (l, r)


def _cone_pairs(n, seed):
    """Trial strains beyond the Drucker-Prager cone and their FP64 projections (oracle dp_project)."""
    rng = np.random.default_rng(seed)
    p = scenes.sand_params()
    lam_over_mu = 2 * 0.3 / (1 - 2 * 0.3)
    kmu = lam_over_mu + 2.0 / 3.0
    inp = np.empty((n, 5), np.float32)
    H = np.empty((n, 3), np.float32)
    c = (3 * lam_over_mu + 2) / 2
    for i in range(n):
        phi = rng.uniform(25.0, 40.0)
        sf = np.sin(np.radians(phi))
        alpha = np.sqrt(2.0 / 3.0) * 2 * sf / (3 - sf)
        tr = -rng.uniform(1e-3, 3e-2)
        d = rng.normal(size=3)
        d -= d.mean()
        d *= c * alpha * abs(tr) * rng.uniform(1.05, 3.0) / np.linalg.norm(d)
        eps = np.sort(tr / 3 + d)[::-1]
        Sn, dq = ob.dp_project(p, phi, 0.0, np.exp(eps))
        assert dq > 0
        inp[i] = (*eps, phi, kmu)
        H[i] = np.log(Sn)
    return inp, H


def test_surrogate_training_deterministic_and_converges():
    inp, H = _cone_pairs(6000, 5)
    runs = []
    for rep in range(2):
        with _solver() as s:
            s.surrogate_init(9)
            say(f"[surrogate train] rep {rep}")
            losses = s.surrogate_train(inp, H, epochs=30, batch=256, lr=2e-3, seed=11)
            runs.append((losses, s.surrogate_get_weights()))
    (l0, w0), (l1, w1) = runs
    say(f"  epoch losses {l0[0]:.4e} -> {l0[-1]:.4e}")
    assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)), "training is not deterministic"
    assert np.array_equal(l0, l1)
    assert l0[-1] <= 0.1 * l0[0]
    assert np.all(w0[~mo.active_mask()] == 0.0)


def _impact_scene():
    """A sand block driven into the ground: most particles take the cone branch of the return map."""
    return scenes.small_scene(_lib.SAND, 2, 0, _lib.FRICTION_COULOMB, seed=3, res=(32, 32, 32), cells=(8, 12, 8),
                              ppa=2, lo=(0.35, 0.19, 0.35), hi=(0.6, 0.5, 0.6), v0=(0.0, -2.0, 0.0), dt=1e-4)


def test_pair_logging_matches_oracle():
    sc = _impact_scene()
    with _solver() as s:
        sc.load_into(s)
        s.set_pair_logging(400000)
        say(f"[pair logging] n={sc.n}")
        s.run(sc.dt, 30)
        st = s.stats()
        pairs = s.get_pairs()
    say(f"  pairs logged {st['pairs_logged']}")
    assert st["pairs_logged"] > 1000 and pairs.shape == (st["pairs_logged"], 8)
    p = sc.params
    lam_over_mu = 2 * p.poisson_ratio / (1 - 2 * p.poisson_ratio)
    assert np.allclose(pairs[:, 4], lam_over_mu + 2.0 / 3.0, rtol=1e-6)
    assert np.all(pairs[:, 3] == np.float32(p.dp_h[0]))
    assert np.all(pairs[:, 0] >= pairs[:, 1]) and np.all(pairs[:, 1] >= pairs[:, 2])
    idx = np.random.default_rng(0).choice(len(pairs), min(3000, len(pairs)), replace=False)
    worst = 0.0
    for i in idx:
        Sn, dq = ob.dp_project(p, float(pairs[i, 3]), 0.0, np.exp(pairs[i, :3].astype(np.float64)))
        worst = max(worst, float(np.max(np.abs(np.log(Sn) - pairs[i, 5:8]))))
    say(f"  max |H - oracle| = {worst:.3e}")
    assert worst <= 2e-6


def _run(sc, nsteps, surrogate=None, gate=1e-5, weights=None):
    with _solver() as s:
        sc.load_into(s)
        if surrogate is not None:
            s.surrogate_set_weights(weights)
            s.set_surrogate(True, gate)
        s.run(sc.dt, nsteps)
        out = s.get_particles()
        out["stats"] = s.stats()
    return out


def test_surrogate_all_rejected_is_bitwise_exact():
    sc = _impact_scene()
    w = np.zeros(mo.NPARAM, np.float32)
    w[mo.active_mask()] = np.nan
    say(f"[surrogate NaN network] n={sc.n}")
    ref = _run(sc, 20)
    sur = _run(sc, 20, surrogate=True, weights=w)
    st = sur["stats"]
    say(f"  candidates {st['surrogate_candidates']} accepted {st['surrogate_accepted']}")
    assert st["surrogate_candidates"] > 1000 and st["surrogate_accepted"] == 0
    for k in ("x", "v", "B", "F", "plastic"):
        assert np.array_equal(ref[k].view(np.uint32), sur[k].view(np.uint32)), k


def test_surrogate_trained_is_accepted_and_close():
    sc = _impact_scene()
    nsteps = 30
    with _solver() as s:
        sc.load_into(s)
        s.set_pair_logging(1 << 20)
        s.run(sc.dt, nsteps)
        pairs = s.get_pairs()
        pairs = pairs[np.lexsort(pairs.T[::-1])]  # append order varies between runs
        say(f"[surrogate trained] pairs {len(pairs)}")
        s.surrogate_init(1)
        losses = s.surrogate_train(pairs[:, :5], pairs[:, 5:8], epochs=60, batch=512, lr=2e-3, seed=3)
        w = s.surrogate_get_weights()
    say(f"  loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    ref = _run(sc, nsteps)
    sur = _run(sc, nsteps, surrogate=True, gate=2e-5, weights=w)
    st = sur["stats"]
    rate = st["surrogate_accepted"] / max(1, st["surrogate_candidates"])
    dx = float(np.max(np.abs(sur["x"] - ref["x"])))
    say(f"  candidates {st['surrogate_candidates']} accepted {st['surrogate_accepted']} ({rate:.3f}); max|dx| {dx:.3e}")
    assert st["surrogate_accepted"] > 0
    assert dx <= 1e-4
    assert st["status"] == 0


def _trained_weights(sc, nsteps=30, hold_out=0.0, seed=3):
    """Pairs logged by the exact path over `nsteps` substeps of sc; the network trained on the GPU on
    all of them, or on a random (1 - hold_out) share. Returns weights, training and held-out pairs."""
    with _solver() as s:
        sc.load_into(s)
        s.set_pair_logging(1 << 20)
        s.run(sc.dt, nsteps)
        pairs = s.get_pairs()
    # the log's row order is the order in which threads appended to it (an atomic counter): sorted
    # rows make the split and the training data the same in every run
    pairs = pairs[np.lexsort(pairs.T[::-1])]
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(pairs))
    k = int(round(len(pairs) * (1.0 - hold_out)))
    train, held = pairs[idx[:k]], pairs[idx[k:]]
    with _solver() as s:
        s.surrogate_init(1)
        losses = s.surrogate_train(train[:, :5], train[:, 5:8], epochs=300, batch=512, lr=2e-3, seed=seed)
        w = s.surrogate_get_weights()
    say(f"  pairs {len(pairs)} (trained on {len(train)}), loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    return w, train, held


def test_gate_decisions_match_oracle():
    """The device's accept flags against the oracle's gate (oracle/mlp_oracle.gate, the yield-residual,
    volume and flow checks in float64) on the device's own trial strains and predictions, recorded for
    every candidate of one substep: equal wherever no gate quantity lies within 1e-8 (1e-3 tol, about
    ten FP32 roundings of strains of 1e-2) of its threshold. The recorded predictions H against the
    oracle network (BF16 roundings emulated) on the same inputs: |dH| <= 1e-3 (2e-3 max(1, |y|) + 1e-3),
    the forward test's bar in H units."""
    sc = _impact_scene()
    tol = 2e-5
    say(f"[gate decisions] n={sc.n}")
    w, _, _ = _trained_weights(sc)
    with _solver() as s:
        sc.load_into(s)
        s.surrogate_set_weights(w)
        s.set_surrogate(True, tol)
        s.set_surrogate_log(1 << 20)
        s.run(sc.dt, 29)
        s.substep(sc.dt)
        rec = s.get_surrogate_log()
        st = s.stats()
    m = len(rec["slot"])
    say(f"  candidates in the last substep {m}, accepted {int(rec['accepted'].sum())}")
    assert m > 200 and 0 < rec["accepted"].sum() < m
    p = sc.params
    lam_over_mu = 2 * p.poisson_ratio / (1 - 2 * p.poisson_ratio)
    c = (3 * lam_over_mu + 2) / 2
    assert np.allclose(rec["kmu"], lam_over_mu + 2.0 / 3.0, rtol=1e-6)
    acc, mg = mo.gate(rec["eps"], rec["H"], rec["phi"], c, tol)
    tie = np.min(np.abs(mg), axis=1) < 1e-8
    agree = acc == rec["accepted"]
    say(f"  oracle accepts {int(acc.sum())}; disagreements {int((~agree).sum())}, all within ties: "
        f"{bool(np.all(tie[~agree]))}; ties {int(tie.sum())}")
    assert np.all(agree | tie)
    assert tie.sum() <= 0.01 * m
    inp = np.concatenate([rec["eps"], rec["phi"][:, None], rec["kmu"][:, None]], axis=1)
    y_ref, _ = mo.forward(w, inp, emulate_bf16=True)
    H_ref = inp[:, :3].astype(np.float64) + mo.CORR_SCALE * y_ref
    dH = np.max(np.abs(rec["H"] - H_ref))
    say(f"  max |H - oracle network| = {dH:.3e}")
    assert dH <= 1e-3 * (2e-3 * max(1.0, np.max(np.abs(y_ref))) + 1e-3)
    # the slots are distinct sorted positions of sand particles
    assert len(np.unique(rec["slot"])) == m and np.all((rec["slot"] >= 0) & (rec["slot"] < sc.n))
    assert st["status"] == 0


def test_trained_network_gains():
    """What the trained network gains over the method without it, on held-out (trial, projected) pairs
    of the impact scene (20% of the logged pairs, not trained on): the gate (oracle/mlp_oracle.gate at
    2e-5) accepts at least half of the trained network's predictions, at most a tenth as many of the
    untrained network's (it passes only trial states within the tolerance of the cone) and none of
    the prediction without a network (H = eps, no plastic correction); the trained network's mean
    error |H - H_exact| is at most half the no-network error (the correction itself) and half the
    untrained network's."""
    sc = _impact_scene()
    tol = 2e-5
    say(f"[trained network gains] n={sc.n}")
    w, _, held = _trained_weights(sc, hold_out=0.2)
    p = sc.params
    lam_over_mu = 2 * p.poisson_ratio / (1 - 2 * p.poisson_ratio)
    c = (3 * lam_over_mu + 2) / 2
    inp, H_exact = held[:, :5], held[:, 5:8].astype(np.float64)
    with _solver() as s:
        s.surrogate_set_weights(w)
        H_tr = s.surrogate_predict(inp).astype(np.float64)
        s.surrogate_init(1)
        H_un = s.surrogate_predict(inp).astype(np.float64)
    corr = np.max(np.abs(H_exact - inp[:, :3]), axis=1)
    out = {}
    for name, H in (("trained", H_tr), ("untrained", H_un), ("no network", inp[:, :3].astype(np.float64))):
        acc, _ = mo.gate(inp[:, :3], H, inp[:, 3], c, tol)
        err = np.max(np.abs(H - H_exact), axis=1)
        out[name] = (acc.mean(), err.mean(), np.median(err / corr))
        say(f"  {name:10s}: accepted {acc.mean():.3f}, mean |H - H_exact| {err.mean():.3e}, "
            f"median |H - H_exact| / |correction| {out[name][2]:.3e}")
    assert out["trained"][0] >= 0.5
    assert out["untrained"][0] <= 0.1 * out["trained"][0] and out["no network"][0] == 0.0
    assert out["trained"][1] <= 0.5 * out["no network"][1] and out["trained"][1] <= 0.5 * out["untrained"][1]
