"""Differentiable mean-field relaxation of the SCA dynamics with the Onsager reaction term (build-defined path).

Dynamics per instance b (the expectation of the discrete decision of fpga/v80_sca/src/sca_ref_bias.hpp:53-56 with tanh in
place of the uniform-noise clip, and sum_i (1 - m_i^2) in place of the band count n_lin):

    u_t = J m_t + q m_t - beta_t m_{t-1},  m_{t+1} = tanh(u_t / (2 T_t))
    temperature-scaled: beta_t = kappa * gamma * T_{t-1} * rho_t          (kconst of mb::tables, mb_common.hpp:149-153)
    counted:            beta_t = lambda * rho_t * gamma / (2 T_{t-1}) * sum_i (1 - m_{t,i}^2)   (kcorr * n_lin)

The schedule T_t is piecewise linear in 16 knots; the forward and the adjoint (with recomputation from checkpoints every
`ckpt` steps) run in HIP kernels (anc_mf_forward / anc_mf_backward); torch autograd only chains T(knots) and
beta(kappa or lambda, T). The tuned schedule is judged by the bit-exact discrete engine (``discrete_schedule``).
"""
import ctypes

import numpy as np
import torch

from . import _native as N


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ramp(S, on=True, device=None, dtype=torch.float32):
    """rho_t: 1 for t < 0.7 S, then (S-1-t)/max(1, S-1-0.7 S) (mb_common.hpp:146-153)."""
    t = torch.arange(S, device=device, dtype=torch.float64)
    if not on:
        return torch.ones(S, device=device, dtype=dtype)
    r = torch.where(t >= 0.7 * S, (S - 1 - t) / max(1.0, S - 1 - 0.7 * S), torch.ones_like(t))
    return r.to(dtype)


def knots_to_T(theta, S):
    """Piecewise-linear temperatures over S steps from K knots at t_k = k (S-1)/(K-1) (differentiable in theta)."""
    K = theta.shape[0]
    pos = torch.arange(S, device=theta.device, dtype=theta.dtype) * (K - 1) / max(1, S - 1)
    i0 = torch.clamp(pos.floor().long(), 0, K - 2)
    w = pos - i0.to(theta.dtype)
    return theta[i0] * (1 - w) + theta[i0 + 1] * w


def coefficients(T, strength, mode, gamma=1.0, ramp_on=True):
    """coef_t (t >= 1) for the two correction modes; coef_0 = 0."""
    S = T.shape[0]
    rho = ramp(S, ramp_on, T.device, T.dtype)
    prevT = torch.cat([T[:1], T[:-1]])
    if mode == "kappa":
        c = strength * gamma * prevT * rho
    elif mode == "counted":
        c = strength * rho * gamma / (2.0 * prevT)
    else:
        raise ValueError(mode)
    return torch.cat([torch.zeros_like(c[:1]), c[1:]])


class Problem:
    """B instances of n spins on the device. jbits = 8 keeps J as int8 rows (row stride ld, a multiple of 16); jbits
    = 4 or 2 keep the packed copy of anc_mf_pack_j4 (entries in [-8, 7]) or anc_mf_pack_j2 ([-2, 1], e.g. +-1
    couplings): a half or a quarter of the bytes per step. The default (None) takes the smallest that fits. All
    storages give identical floats."""

    RANGES = {2: (-2, 1), 4: (-8, 7), 8: (-128, 127)}

    def __init__(self, J, device="cuda", jbits=None):
        J = torch.as_tensor(J)
        if J.dim() == 2:
            J = J.unsqueeze(0)
        B, n, _ = J.shape
        J = J[:, :n, :n]
        lo, hi = (int(J.min()), int(J.max())) if J.numel() else (0, 0)
        fits = [b for b in (2, 4, 8) if self.RANGES[b][0] <= lo and hi <= self.RANGES[b][1]]
        if jbits is None:
            jbits = fits[0]
        if jbits not in (2, 4, 8):
            raise ValueError("jbits must be 2, 4 or 8")
        if jbits not in fits:
            raise ValueError(f"jbits = {jbits} needs every coupling in {list(self.RANGES[jbits])}")
        ld = ((n + 15) // 16) * 16
        Jp = torch.zeros((B, n, ld), dtype=torch.int8, device=device)
        Jp[:, :, :n] = J.to(device=device, dtype=torch.int8)
        self.B, self.n, self.jbits = B, n, jbits
        if jbits == 8:
            self.J, self.ld = Jp.contiguous(), ld
        else:
            ldp = ((2 * jbits * ((n + 15) // 16) + 15) // 16) * 16
            Jq = torch.empty((B, n, ldp), dtype=torch.uint8, device=device)
            bad = torch.zeros(1, dtype=torch.int32, device=device)
            pack = N.lib().anc_mf_pack_j4 if jbits == 4 else N.lib().anc_mf_pack_j2
            N.check(pack(B, n, _p(Jp), ld, n * ld, _p(Jq), ldp, n * ldp, _p(bad), _stream()))
            if int(bad.item()):
                raise RuntimeError(f"packing found couplings outside {list(self.RANGES[jbits])}")
            self.J, self.ld = Jq, ldp
        self.bstride = n * self.ld

    @property
    def j_bytes(self):
        """Bytes of J one GEMV pass over all instances streams."""
        return self.B * self.bstride

    def gemv(self, x):
        y = torch.empty_like(x)
        N.check(N.lib().anc_mf_gemv(self.B, self.n, _p(self.J), self.jbits, self.ld, self.bstride, _p(x), _p(y), _stream()))
        return y


class _Relax(torch.autograd.Function):
    @staticmethod
    def forward(ctx, T, coef, m0, prob, q, counted, ckpt):
        ctx.t_dtype, ctx.c_dtype = T.dtype, coef.dtype
        T = T.detach().float().contiguous()
        coef = coef.detach().float().contiguous()
        m0 = m0.detach().float().contiguous()
        S = T.shape[0]
        B, n = prob.B, prob.n
        nseg = (S + ckpt - 1) // ckpt
        dev = m0.device
        cbuf = torch.empty(2 * nseg * B * n, dtype=torch.float32, device=dev)
        mS = torch.empty((B, n), dtype=torch.float32, device=dev)
        work = torch.empty(N.lib().anc_mf_workspace(B, n, ckpt) // 4 + 16, dtype=torch.float32, device=dev)
        N.check(N.lib().anc_mf_forward(B, n, _p(prob.J), prob.jbits, prob.ld, prob.bstride, ctypes.c_float(q), S, _p(T), _p(coef), int(counted),
                                       _p(m0), ckpt, _p(cbuf), _p(mS), _p(work), _stream()))
        N.check(N.lib().anc_mf_status(_p(work), B, n, ckpt, _stream()))
        ctx.save_for_backward(T, coef, cbuf)
        ctx.prob, ctx.q, ctx.counted, ctx.ckpt, ctx.work = prob, q, counted, ckpt, work
        ctx.m0_grad = m0.requires_grad
        return mS

    @staticmethod
    def backward(ctx, gS):
        T, coef, cbuf = ctx.saved_tensors
        prob = ctx.prob
        S, B, n = T.shape[0], prob.B, prob.n
        gS = gS.detach().float().contiguous()
        dT = torch.empty((S, B), dtype=torch.float64, device=gS.device)
        dc = torch.empty((S, B), dtype=torch.float64, device=gS.device)
        g0 = torch.empty((B, n), dtype=torch.float32, device=gS.device)
        N.check(N.lib().anc_mf_backward(B, n, _p(prob.J), prob.jbits, prob.ld, prob.bstride, ctypes.c_float(ctx.q), S, _p(T), _p(coef),
                                        int(ctx.counted), ctx.ckpt, _p(cbuf), _p(gS), _p(dT), _p(dc), _p(g0), _p(ctx.work), _stream()))
        N.check(N.lib().anc_mf_status(_p(ctx.work), B, n, ctx.ckpt, _stream()))
        return dT.sum(1).to(ctx.t_dtype), dc.sum(1).to(ctx.c_dtype), g0, None, None, None, None


def set_team(on):
    """Process-wide choice of the mean-field kernels: persistent team kernels where they apply (n <= 2048, the
    default) or one launch per step (results are identical). Returns the previous setting."""
    old = bool(N.lib().anc_mf_get_option(b"team"))
    N.check(N.lib().anc_mf_set_option(b"team", 1.0 if on else 0.0))
    return old


def last_team():
    """Whether the last forward or backward pass ran the persistent team kernels."""
    return bool(N.lib().anc_mf_get_option(b"last_team") == 1)


def relax(prob, T, coef, m0, q, counted=False, ckpt=32):
    """m_S of every instance (B, n); differentiable in T, coef (and m0)."""
    return _Relax.apply(T, coef, m0, prob, float(q), bool(counted), int(ckpt))


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, prob):
        m = m.detach().float().contiguous()
        h = prob.gemv(m)
        ctx.save_for_backward(h)
        return -0.5 * (m * h).sum(1)

    @staticmethod
    def backward(ctx, g):
        (h,) = ctx.saved_tensors
        return -g[:, None] * h, None


def energy(prob, m):
    """E_b = -1/2 m_b . J_b m_b for every instance (differentiable in m)."""
    return _Energy.apply(m, prob)


def loss(prob, theta, strength, m0, q, S, mode="kappa", gamma=1.0, ckpt=32):
    """Mean energy per spin after S relaxation steps with the schedule from knots theta and the correction strength."""
    # the schedule algebra (knots -> T_t, kappa or lambda -> coef_t) and its chain rule run in FP64
    T = knots_to_T(theta.double(), S)
    coef = coefficients(T, strength.double() if torch.is_tensor(strength) else strength, mode, gamma)
    mS = relax(prob, T, coef, m0, q, counted=(mode == "counted"), ckpt=ckpt)
    return energy(prob, mS).double().mean() / prob.n


def reference_relax(J, T, coef, m0, q, counted=False):
    """FP64 torch reference of the same dynamics (autograd through every step). J: (B, n, n) float64."""
    m_prev, m = m0, m0
    S = T.shape[0]
    for t in range(S):
        beta = coef[t] * (1 - m * m).sum(1, keepdim=True) if counted else coef[t]
        if t == 0:
            beta = beta * 0
        u = torch.bmm(J, m.unsqueeze(2)).squeeze(2) + q * m - beta * m_prev
        m_prev, m = m, torch.tanh(u / (2 * T[t]))
    return m


def reference_loss(J, theta, strength, m0, q, S, mode="kappa", gamma=1.0):
    T = knots_to_T(theta, S)
    coef = coefficients(T, strength, mode, gamma)
    mS = reference_relax(J, T, coef, m0, q, counted=(mode == "counted"))
    e = -0.5 * (mS * torch.bmm(J, mS.unsqueeze(2)).squeeze(2)).sum(1)
    return e.mean() / J.shape[1]


def discrete_schedule(theta, strength, S, mode="kappa", q=8.0, gamma=1.0):
    """The tuned schedule as an explicit-T schedule of the discrete engine (same kappa / lambda rule, scale = gamma)."""
    T = knots_to_T(theta.detach().double().cpu(), S).numpy()
    s = dict(S=S, shape="explicit", T=T, q=q, ramp=True, scale=gamma)
    if mode == "kappa":
        s["kappa"] = float(strength)
    else:
        s["lam"] = float(strength)
    return s
