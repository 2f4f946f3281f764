"""Differentiable astrometry of N-body orbits on the GPU (PyTorch-ROCm).

``radec(params, ...)`` maps fit parameters to predicted RA/Dec at the
observation epochs; its backward pass multiplies the incoming gradient by the
partial derivatives that the forward launch computes alongside the values
(``kh_radec_partials_dev``, forward-mode sensitivities of the Radau
integrator, the light-time step and the RA/Dec projection).  ``chi2`` builds
the weighted residual objective of an orbit fit on top of it.

kete has no derivatives of its propagators; the observable chain is the one
its FOV check uses (``fov/fov_like.rs:89-110``): ``propagate_n_body_spk`` to
each epoch, one two-body step back by the light time, RA/Dec of the
observer-to-object vector (``simult_states.rs:86-120``).

Parameters per object, in this order: x, y, z (AU), vx, vy, vz (AU/day),
Sun-centred in the ecliptic (or equatorial) frame at ``jd0``, and A1, A2, A3
(AU/day^2) of the JplComet model with kete's ``new_asteroid`` defaults
(``src/kete/rust/nongrav.rs:128-161``: alpha 1, r_0 1, m 2, n 1, k 0).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import FRAME_ECLIPTIC, _lib, get_context

N_PARAMS = 9


def _asteroid_models(a123: torch.Tensor) -> torch.Tensor:
    """kh_nongrav records (88 bytes: int32 kind, pad, 10 doubles) on the
    device for JplComet with the new_asteroid defaults and the given A."""
    n = a123.shape[0]
    rec = torch.zeros((n, 11), dtype=torch.float64, device=a123.device)
    rec.view(torch.int64)[:, 0] = 1          # kind 1 (JplComet), pad 0
    rec[:, 1:4] = a123
    rec[:, 4] = 1.0                          # alpha
    rec[:, 5] = 1.0                          # r_0
    rec[:, 6] = 2.0                          # m
    rec[:, 7] = 1.0                          # n
    rec[:, 8] = 0.0                          # k
    return rec.contiguous()


def _library_stream(ctx):
    """The context's HIP stream as a torch stream object."""
    s = getattr(ctx, "_torch_stream", None)
    if s is None:
        s = torch.cuda.ExternalStream(ctx.L.kh_ctx_stream(ctx.h), device=f"cuda:{ctx.device}")
        ctx._torch_stream = s
    return s


def radec_partials_dev(params, jd0, jd_obs, obs, frame=FRAME_ECLIPTIC, force=0, epsilon=1e-6,
                       ctx=None, counters=None, restart=False):
    """One launch on device tensors.

    params: [n, 9] float64 (cuda); jd0: [n]; jd_obs: [m, n]; obs: [3, m, n];
    counters: optional int32 [n, 4] device tensor receiving (force
    evaluations, sweeps, accepted steps, failed steps) per object.
    -> (radec [2, m, n], jac [2, 9, m, n], status [n] int32)."""
    ctx = ctx or get_context()
    dev = params.device
    n = params.shape[0]
    m = jd_obs.shape[0]
    cols = [params[:, k].contiguous() for k in range(6)]
    ng = _asteroid_models(params[:, 6:9].contiguous())
    jd0 = jd0.to(dev, torch.float64).contiguous()
    jd_obs = jd_obs.to(dev, torch.float64).contiguous()
    obs = obs.to(dev, torch.float64).contiguous()
    radec = torch.empty((2, m, n), dtype=torch.float64, device=dev)
    jac = torch.empty((2, 9, m, n), dtype=torch.float64, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    # the library's stream waits for torch's work on the inputs, and torch's
    # stream for the library's results: stream order, no host round trip
    lib_stream = _library_stream(ctx)
    cur = torch.cuda.current_stream(dev)
    lib_stream.wait_stream(cur)
    _lib.check(ctx.L.kh_radec_partials_dev(
        ctx.h, n, _lib.raw(jd0.data_ptr()), _lib.ptr_array([c.data_ptr() for c in cols]), frame,
        _lib.raw(ng.data_ptr(), _lib.KhNonGrav), m, _lib.raw(jd_obs.data_ptr()),
        _lib.ptr_array([obs[k].data_ptr() for k in range(3)]), force, int(restart), epsilon,
        _lib.ptr_array([radec[0].data_ptr(), radec[1].data_ptr()]),
        _lib.ptr_array([jac[0].data_ptr(), jac[1].data_ptr()]), None, None,
        _lib.raw(status.data_ptr(), _lib.C.c_int32),
        _lib.raw(counters.data_ptr(), _lib.KhCounters) if counters is not None else None, 0))
    cur.wait_stream(lib_stream)
    for t in (radec, jac, status) + ((counters,) if counters is not None else ()):
        t.record_stream(lib_stream)
    return radec, jac, status


class RaDec(torch.autograd.Function):
    """params [n, 9] -> predicted (RA, Dec) [n, m, 2]; backward applies the
    partials computed in the same launch."""

    @staticmethod
    def forward(fctx, params, jd0, jd_obs, obs, frame, force, epsilon, restart):
        radec, jac, status = radec_partials_dev(params.detach(), jd0, jd_obs, obs, frame, force,
                                                epsilon, restart=restart)
        fctx.save_for_backward(jac)
        fctx.status = status
        return radec.permute(2, 1, 0).contiguous()

    @staticmethod
    def backward(fctx, grad):
        (jac,) = fctx.saved_tensors
        g = grad.permute(2, 1, 0).contiguous()           # [2, m, n]
        gp = (jac * g[:, None]).sum(dim=(0, 2))          # [9, n]
        return gp.t().contiguous(), None, None, None, None, None, None, None


def radec(params, jd0, jd_obs, obs, frame=FRAME_ECLIPTIC, force=0, epsilon=1e-6, restart=False):
    """Differentiable predicted RA/Dec [n, m, 2] (radians).  restart: a new
    integrator at each epoch (as successive propagate_n_body calls) instead of
    one integration through all epochs; the two agree to the integrator's
    error level."""
    return RaDec.apply(params, jd0, jd_obs, obs, frame, force, epsilon, restart)


def chi2(params, jd0, jd_obs, obs, ra_obs, dec_obs, sigma, frame=FRAME_ECLIPTIC, force=0,
         epsilon=1e-6, restart=False):
    """Sum over objects and epochs of ((dRA cos Dec_obs)^2 + dDec^2) / sigma^2,
    with dRA wrapped to [-pi, pi).  ra_obs, dec_obs: [n, m]; sigma in radians
    (scalar or [n, m]).  Returns a scalar tensor."""
    p = radec(params, jd0, jd_obs, obs, frame, force, epsilon, restart)
    dra = torch.remainder(p[..., 0] - ra_obs + math.pi, 2 * math.pi) - math.pi
    ddec = p[..., 1] - dec_obs
    return (((dra * torch.cos(dec_obs)) ** 2 + ddec ** 2) / sigma ** 2).sum()


_OBL_C, _OBL_S = math.cos(0.40909280422232897), math.sin(0.40909280422232897)


def radec_plain(params, jd0, jd_obs, obs, frame=FRAME_ECLIPTIC, epsilon=1e-6, ctx=None):
    """The same observables without derivatives, from the plain propagators:
    per epoch one kh_propagate_n_body_dev launch (all objects to the epoch),
    one kh_propagate_two_body_dev launch with the observer (the light-time
    step, kepler.rs:68-108), then RA/Dec on the device.  Epochs are shared by
    all objects here.  Used as the finite-difference baseline.

    params: [N, 9] (cuda); jd0: [N]; jd_obs: [m] (host); obs: [3, m] (host,
    observer w.r.t. the Sun in `frame`).  -> (radec [N, m, 2], status [N])."""
    ctx = ctx or get_context()
    L = ctx.L
    dev = params.device
    n = params.shape[0]
    jd_obs = np.asarray(jd_obs, dtype=np.float64)
    obs = np.ascontiguousarray(obs, dtype=np.float64)
    m = len(jd_obs)
    state = params[:, :6].t().contiguous()
    ng = _asteroid_models(params[:, 6:9].contiguous())
    t = jd0.to(dev, torch.float64).contiguous()
    epochs = torch.as_tensor(jd_obs, dtype=torch.float64, device=dev)
    t_vecs = [torch.full((n,), float(jd_obs[j]), dtype=torch.float64, device=dev)
              for j in range(m)]
    out = torch.empty((n, m, 2), dtype=torch.float64, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    st_nb = torch.empty(n, dtype=torch.int32, device=dev)
    st_tb = torch.empty(n, dtype=torch.int32, device=dev)
    new = torch.empty_like(state)
    lt = torch.empty_like(state)
    cols = lambda s: _lib.ptr_array([s[k].data_ptr() for k in range(6)])
    i32 = lambda x: _lib.raw(x.data_ptr(), _lib.C.c_int32)
    for j in range(m):
        torch.cuda.synchronize(dev)
        _lib.check(L.kh_propagate_n_body_dev(
            ctx.h, n, _lib.raw(t.data_ptr()), cols(state), 10, frame,
            _lib.raw(epochs[j:].data_ptr()), 0, 0, _lib.raw(ng.data_ptr(), _lib.KhNonGrav),
            epsilon, -1, cols(new), i32(st_nb), None, None, None, 0))
        ob = np.ascontiguousarray(obs[:, j])
        _lib.check(L.kh_propagate_two_body_dev(
            ctx.h, n, _lib.raw(t_vecs[j].data_ptr()), cols(new), 10, frame, 1,
            _lib.raw(epochs[j:].data_ptr()), _lib.dptr(ob), cols(lt), i32(st_tb), None, 1))
        status = torch.where(status == 0, st_nb, status)
        status = torch.where(status == 0, st_tb, status)
        rho = lt[:3] - torch.as_tensor(ob, dtype=torch.float64, device=dev)[:, None]
        if frame == FRAME_ECLIPTIC:
            rho = torch.stack([rho[0], _OBL_C * rho[1] - _OBL_S * rho[2],
                               _OBL_S * rho[1] + _OBL_C * rho[2]])
        r = torch.linalg.vector_norm(rho, dim=0)
        out[:, j, 0] = torch.remainder(torch.atan2(rho[1], rho[0]), 2 * math.pi)
        out[:, j, 1] = torch.asin(torch.clamp(rho[2] / r, -1.0, 1.0))
        state, new = new, state
        t = t_vecs[j]
    return out, status


def synthetic_observations(n=4096, n_obs=60, years=2.0, noise_arcsec=0.1, seed_index=5,
                           ctx=None, device="cuda:0"):
    """Seeded fitting workload: NEO-like orbits with A1..A3 ~ N(0, 1e-12),
    observation epochs uniform over `years`, observer the ephemeris's Earth,
    observations from the true orbits plus Gaussian RA/Dec noise.
    -> dict of device tensors (params_true, jd0, jd_obs, obs, ra_obs, dec_obs, sigma)."""
    from . import synthetic as syn
    ctx = ctx or get_context()
    rng = np.random.default_rng(syn.SEED_BASE + seed_index)
    xyz = syn.n_body_workload(n=n, neo_fraction=1.0, seed_index=seed_index)
    a123 = rng.normal(0.0, 1e-12, (n, 3))
    jd0 = syn.EPOCH_JD
    jd_obs = np.sort(jd0 + rng.uniform(0.0, years * 365.25, n_obs))
    earth, st = ctx.spk_state(399, jd_obs, 10, FRAME_ECLIPTIC)
    if np.any(st != 0):
        raise RuntimeError("observer epochs outside the ephemeris")
    params = np.concatenate([xyz.T, a123], axis=1)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    jd0_t = t(np.full(n, jd0))
    jo = t(np.broadcast_to(jd_obs[:, None], (n_obs, n)))
    ob = t(np.broadcast_to(earth[:3, :, None], (3, n_obs, n)))
    pt = t(params)
    with torch.no_grad():
        truth = radec(pt, jd0_t, jo, ob)
    sigma = noise_arcsec / 206264.80624709636
    noise = t(rng.normal(0.0, sigma, (n, n_obs, 2)))
    return dict(params_true=pt, jd0=jd0_t, jd_obs=jo, obs=ob,
                ra_obs=torch.remainder(truth[..., 0] + noise[..., 0], 2 * math.pi),
                dec_obs=truth[..., 1] + noise[..., 1], sigma=sigma)
