"""Inverse problem: recover per-tissue mu_a and mu_s from a measured
transmitted image by gradient descent in log-parameters, as the reference's
driver does (FisherInfo_IO/src/inverse_rte.cu:202-368):

    residual   = sum_j (log E_j - log g_j)^2                      (:361-364)
    grad_a[t]  = -2 mu_a[t] sum_j (log g_j - log E_j) dE_j/dmu_a[t] / E_j   (:255)
    log(mu_a[t] / av_a) <- log(mu_a[t] / av_a) - step_a grad_a[t]  (:332-333)

and the same for mu_s (:295, :336-337); a parameter whose gradient magnitude
falls below the threshold is frozen from then on (:323-326).  av_a = 0.2 and
av_s = 1 in the reference (:115-116).  Pixels with E_j <= 0 or g_j <= 0 are
left out of the sums (their logarithm is undefined).

dE/dmu[t] comes from one forward and one adjoint series per iteration
(rte_exitance_vjp, with the cotangent of the residual) summed over the voxels
of tissue t, or, with method="reference", from the reference's own
forward-mode derivative series, one per tissue type (rte_ref_jacobian).
"""
import numpy as np

from . import rte


def reconstruct(ctx, shape, bounds, tissue, mua, mus, g, beam, measured, params, iterations=10, step_a=1e3,
                step_s=1e3, thresh_a=1e-7, thresh_s=1e-7, av_a=0.2, av_s=1.0, n=1.0, method="adjoint",
                fit_a=True, fit_s=True):
    """tissue: uint8 map (x fastest) with types 1..T; mua, mus: length T+1 arrays
    (entry 0 unused).  Returns (mua, mus, history) with history[k] = dict(residual,
    grad_a, grad_s, mua, mus) before the k-th update."""
    tissue = np.asarray(tissue).ravel()
    mua = np.array(mua, dtype=np.float64)
    mus = np.array(mus, dtype=np.float64)
    T = len(mua) - 1
    g_meas = np.asarray(measured, dtype=np.float64).ravel()
    part = tissue.astype(np.int32) - 1
    frozen_a = np.zeros(T + 1, dtype=bool)
    frozen_s = np.zeros(T + 1, dtype=bool)
    frozen_a[0] = frozen_s[0] = True
    history = []
    for _ in range(iterations):
        med = ctx.medium(shape, bounds, mua[tissue], mus[tissue], g, n=n)
        try:
            if method == "reference":
                img0, _ = ctx.solve(shape, bounds, mua[tissue], mus[tissue], g, beam, params, n=n)
                E = img0.ravel()
                Ja, _, _ = ctx.ref_jacobian(med, beam, params, part, T, wrt=0)
                Js, _, _ = ctx.ref_jacobian(med, beam, params, part, T, wrt=1)
            else:
                # the cotangent depends on E: one forward solve, then the adjoint with it
                img0, _ = ctx.solve(shape, bounds, mua[tissue], mus[tissue], g, beam, params, n=n)
                E = img0.ravel()
            ok = (E > 0) & (g_meas > 0)
            lr = np.where(ok, np.log(np.where(ok, E, 1.0)) - np.log(np.where(ok, g_meas, 1.0)), 0.0)
            residual = float(np.sum(lr ** 2))
            # d residual / d E_j
            cot = np.where(ok, 2.0 * lr / np.where(ok, E, 1.0), 0.0)
            if method == "reference":
                da = cot @ Ja
                ds = cot @ Js
            else:
                _, ga, gs, _ = ctx.exitance_vjp(med, beam, params, cot)
                da = np.bincount(part[part >= 0], weights=ga[part >= 0], minlength=T)
                ds = np.bincount(part[part >= 0], weights=gs[part >= 0], minlength=T)
        finally:
            med.close()
        grad_a = np.concatenate([[0.0], mua[1:] * da])
        grad_s = np.concatenate([[0.0], mus[1:] * ds])
        history.append(dict(residual=residual, grad_a=grad_a.copy(), grad_s=grad_s.copy(), mua=mua.copy(),
                            mus=mus.copy()))
        frozen_a |= np.abs(grad_a) < thresh_a
        frozen_s |= np.abs(grad_s) < thresh_s
        if fit_a:
            upd = ~frozen_a
            mua[upd] = av_a * np.exp(np.log(mua[upd] / av_a) - step_a * grad_a[upd])
        if fit_s:
            upd = ~frozen_s
            mus[upd] = av_s * np.exp(np.log(mus[upd] / av_s) - step_s * grad_s[upd])
    return mua, mus, history
