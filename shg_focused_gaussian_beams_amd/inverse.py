"""Inverse design of the pump phase: gradient ascent of the SH efficiency over a
per-pixel phase mask phi(x, y) applied to the z = 0 pump, A = A0 exp(i phi), with the
gradient from the library's discrete adjoint (on a THERMAL context the temperature field
is held fixed during the ascent).

d eta / d phi = Re(conj(g) i A), with g = d eta / d A (dJ = Re sum conj(g) dA).
"""
import numpy as np


def mask_gradient(g, field):
    """d eta / d phi for the field A = A0 exp(i phi) given g = d eta / dA."""
    return np.real(np.conj(g.astype(np.complex128)) * 1j * field.astype(np.complex128))


def ascend_phase_mask(ctx, pump0, mask0, iterations, step=0.02, signal=None, checkpoint_interval=0):
    """Normalised gradient ascent: phi += step * grad / max|grad| per iteration.

    Returns (mask, etas) with etas[k] the efficiency before step k (and the final one
    last).  `ctx` keeps its parameters (and, THERMAL, its temperature field)."""
    pump0 = np.asarray(pump0, np.complex64)
    sh = np.zeros_like(pump0) if signal is None else np.asarray(signal, np.complex64)
    phi = np.asarray(mask0, np.float64).copy()
    etas = []
    for _ in range(iterations):
        field = (pump0 * np.exp(1j * phi)).astype(np.complex64)
        ctx.set_fields(field, sh)
        eta, g = ctx.adjoint_field(checkpoint_interval=checkpoint_interval)
        etas.append(eta)
        d = mask_gradient(g, field)
        m = np.abs(d).max()
        if m > 0:
            phi += step * d / m
    field = (pump0 * np.exp(1j * phi)).astype(np.complex64)
    ctx.set_fields(field, sh)
    ctx.forward()
    etas.append(ctx.powers()[2])
    return phi, etas


def ascend_phase_masks(ctxs, pumps, masks, iterations, step=0.02, workers=None, checkpoint_interval=0):
    """The batched form: instance k ascends its own mask on its own context, all of them
    at once from a pool of host threads (the library calls release the GIL, and each
    context owns its stream, so the instances overlap on the GPU and with each other's
    host work).  Returns [(mask, etas)] in instance order, each as ascend_phase_mask."""
    from concurrent.futures import ThreadPoolExecutor
    if not (len(ctxs) == len(pumps) == len(masks)):
        raise ValueError("ctxs, pumps and masks must have one entry per instance")
    workers = len(ctxs) if workers is None else max(1, int(workers))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        futs = [pool.submit(ascend_phase_mask, c, p, m, iterations, step, None, checkpoint_interval)
                for c, p, m in zip(ctxs, pumps, masks)]
        return [f.result() for f in futs]
