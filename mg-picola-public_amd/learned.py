"""Training and evaluation of the learned Green's-function multiplier.

The multiplier M(|k|, a, theta, s) scales the force Green's function of a
coarse run (Nmesh = Nsample) so that its P(k) follows a plain run on a mesh
three times finer from the same initial conditions (the Gaussian field of the
IC generator does not depend on Nmesh below the particle Nyquist cut).  The
network's inputs are |k|/k_Nyquist of the coarse mesh, a, theta = (log10 f_R0 +
6)/2 for f(R) (-1 without modified gravity) and s, the rms of ln(1 + delta) of
the coarse mesh after the deposit.

Training is a fixed-point iteration, all of it on the GPU through libmgpu:
  1. the fine runs are made once, with P(k) at every force evaluation;
  2. the coarse runs are made with the current network (the initial network
     has a zero output layer, M = 1, which is the plain coarse run), logging a,
     s and P(k) at every force evaluation;
  3. every (step, k bin) gives a sample x = (k/k_Nyq, a, theta, s) with target
     ln M(x) + eta ln(P_fine/P_coarse) of that k bin at the output time
     (clipped to +-eta/2 per iteration) and weight = modes.  The output-time
     ratio is what the metric measures and it integrates the force deficit of
     the whole run: in linear theory a force boost 1 + e from a = 1/20 to 1
     raises P by about exp(3.6 e), so eta = 0.28 would land on the fixed point
     in one step; nonlinear growth needs more, eta = 0.6 by default.  (The same-time ratio is a poor signal: at early times the
     coarse estimator's aliasing dominates it.)  The training runs keep the
     shot noise in P(k), which both runs share;
  4. the network is refitted with full-batch Adam (mgpu_mlp_train) and the
     next iteration repeats 2-4;
  5. the iteration is not monotonic (a refit can overshoot into a coarse run
     whose boosted grid-scale forces heat the particles), so every iterate's
     error is measured, the last one's too, and the best network is kept; the
     error is taken on held-out pairs (seeds never sampled) when they are
     given, else on the training pairs.
The plain fine-mesh run stays the reference of the P(k) error.
"""
import math

import numpy as np

import mgpicola as M


def clone(hp, **over):
    """A copy of a HostParams with fields replaced (strings as str)."""
    c = M.HostParams.from_buffer_copy(hp)
    for k, v in over.items():
        setattr(c, k, v.encode() if isinstance(v, str) else v)
    return c


def theta_of(hp):
    fofr = hp.model in (M.MODELS["FOFR"], M.MODELS["FOFR_LCDM"], M.MODELS["FOFRNU"])
    return (math.log10(hp.fofr0) + 6.0) / 2.0 if fofr and hp.modified_gravity_active and hp.fofr0 > 0 else -1.0


def run_logged(hp, mlp=None):
    """Run hp to the end with the multiplier of mlp (None: plain); returns the
    per-force-evaluation log [(a, s, P, k mean, modes)]."""
    sim = M.Sim(hp)
    try:
        sim.set_learned(mlp, True)
        sim.ic()
        while not sim.finished:
            sim.advance()
        return sim.log()
    finally:
        sim.close()


def samples(fine, coarse, hp_coarse, mlp, kmax, eta=0.3, clip=0.3, wpow=1.0, anchor=0.0):
    """Training samples (x, t, w) of one coarse/fine pair of logs; w = modes**wpow.
    anchor > 0 adds, per step, a sample at k = 0 with target ln M = 0 (the
    coarse mesh is exact at the largest scales) and `anchor` times the mean
    weight of the step's samples."""
    kny = math.pi * hp_coarse.Nmesh / hp_coarse.Box
    th = theta_of(hp_coarse)
    X, R, W = [], [], []
    assert len(fine) == len(coarse), "coarse and fine runs must share the time steps"
    _, _, pc1, kc1, nc1 = coarse[-1]
    pf1 = fine[-1][2]
    ok = (nc1 > 0) & (pc1 > 0) & (pf1 > 0) & (kc1 > 0) & (kc1 < kmax)
    k = kc1[ok]
    corr = np.clip(eta * np.log(pf1[ok] / pc1[ok]), -clip, clip)
    wk = nc1[ok] ** wpow
    anchored = []
    for (a, s, _, _, _), (af, _, _, _, _) in zip(coarse, fine):
        assert abs(a - af) < 1e-12
        X.append(np.stack([k / kny, np.full(k.size, a), np.full(k.size, th), np.full(k.size, s)], axis=1))
        R.append(corr)
        W.append(wk)
        if anchor > 0:
            X.append(np.array([[0.0, a, th, s]]))
            R.append(np.zeros(1))
            W.append(np.array([anchor * wk.mean()]))
            anchored.append(sum(len(x) for x in X) - 1)
    X = np.concatenate(X).astype(np.float32)
    base = mlp.forward(X).astype(np.float64) if mlp is not None else np.zeros(len(X))
    base[anchored] = 0.0  # the anchors' target is ln M = 0 itself, not a correction of the current network
    return X, (base + np.concatenate(R)).astype(np.float32), np.concatenate(W).astype(np.float32)


def pofk_error(fine, coarse, kmax=1.0):
    """max over bins with 0 < k < kmax of |P_coarse/P_fine - 1| at the last force
    evaluation (the output time), and the bins' k."""
    _, _, pc, kc, nc = coarse[-1]
    _, _, pf, _, nf = fine[-1]
    ok = (nc > 0) & (nf > 0) & (kc > 0) & (kc < kmax) & (pf > 0)
    return float(np.max(np.abs(pc[ok] / pf[ok] - 1.0))), kc[ok]


def train(pairs, mlp, iterations=6, iters=1500, lr=3e-3, kmax=None, eta=0.6, eta_decay=1.0, log=None,
          kmax_select=None, wpow=1.0, val_pairs=None, anchor=0.0):
    """pairs: [(hp_coarse, hp_fine)]; returns the training history
    [(iteration, loss, mean P(k) error before the refit)].
    Iteration i uses the gain eta * eta_decay**i.  Samples cover k < kmax
    (default: the coarse Nyquist k) with weights modes**wpow (and k = 0
    anchors of relative weight `anchor`, see samples).  The error that
    picks the kept iterate covers k < kmax_select (default: kmax) and is
    measured on val_pairs (held out: never sampled) when given, else on the
    training pairs."""
    keep = dict(pofk_subtract_shotnoise=0)
    fine_logs = [run_logged(clone(f, **keep)) for _, f in pairs]
    val_logs = [run_logged(clone(f, **keep)) for _, f in val_pairs] if val_pairs else None
    hist = []
    best = (math.inf, None, -1)
    for it in range(iterations + 1):
        Xs, Ts, Ws, errs = [], [], [], []
        e = eta * eta_decay ** it
        for (c, _), fl in zip(pairs, fine_logs):
            cl = run_logged(clone(c, **keep), mlp)
            km = kmax if kmax is not None else math.pi * c.Nmesh / c.Box
            if val_logs is None:
                errs.append(pofk_error(fl, cl, kmax_select if kmax_select is not None else km)[0])
            if it < iterations:
                x, t, w = samples(fl, cl, c, mlp, km, eta=e, clip=0.5 * e, wpow=wpow, anchor=anchor)
                Xs.append(x)
                Ts.append(t)
                Ws.append(w)
        if val_logs is not None:
            for (c, _), fl in zip(val_pairs, val_logs):
                km = kmax if kmax is not None else math.pi * c.Nmesh / c.Box
                errs.append(pofk_error(fl, run_logged(clone(c, **keep), mlp),
                                       kmax_select if kmax_select is not None else km)[0])
        err = float(np.mean(errs))
        if err < best[0]:
            best = (err, mlp.params.copy(), it)
        if it == iterations:
            hist.append((it, None, err))
            break
        loss = mlp.train(np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ws), iters, lr)
        hist.append((it, loss, err))
        if log:
            log(f"iteration {it}: P(k) error before refit {err:.4f}, loss {loss:.3e}")
    # the iteration is not monotonic: keep the network with the smallest error on the pairs
    mlp.params = best[1]
    if log:
        log(f"kept the network of iteration {best[2]} (P(k) error {best[0]:.4f})")
    return hist
