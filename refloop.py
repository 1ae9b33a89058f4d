"""The reference's CDL step sequence run as it is on the GPU: the comparison point for the
batched and served forms ("the plain per-dataset GPU loop").

Not part of the library's compute path (that is libcdlm's fused HIP kernels); this is a
measurement baseline.  It keeps the reference's state (the ~25 arrays of cdl{T,S,Q},
src/SETUP/CDL_setup.jl:70-271), and runs code_update!, dict_update! and update_best!
(src/CDL/train.jl:1-98) statement for statement: batched complex FFTs along l (hipFFT through
torch.fft, the cuFFT plans of CDL_setup.jl:158-163,195-204) and one elementwise or reduction
operation per reference kernel or broadcast (csc_kernels.jl, cdl_kernels.jl), with a device
synchronisation wherever the reference has CUDA.@sync.  Shapes are the reference's
column-major (Lg, M, N) held as torch (N, M, Lg), l fastest.  FP32 / complex64.
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


class ReferenceStructuredCDL:
    """cdl{T,S,Q} state on the GPU built from the same inputs as libcdlm's problem:
    s (N, 4L) one-hot, filter lengths, rho, sigma, lambda, D0 (M, Lg) filters before
    masking, X0 (N, M, Lg) codes before the stride-4 projection (numpy or a torch tensor,
    which may already be on the device)."""

    def __init__(self, s, filter_len, rho, sigma, lam, D0, X0, device="cuda"):
        torch = _torch()
        self.t = torch
        self.dev = torch.device(device)
        s = np.ascontiguousarray(s, np.float32)
        self.N, self.L4 = s.shape
        self.fl = [int(f) for f in filter_len]
        self.M = len(self.fl)
        self.gamma = max(self.fl)
        self.Lg = self.L4 + 2 * self.gamma
        self.rho, self.sigma, self.lam = float(rho), float(sigma), float(lam)
        N, M, Lg, g = self.N, self.M, self.Lg, self.gamma
        f32, c64 = torch.float32, torch.complex64
        T = lambda a: (a.to(self.dev, f32) if torch.is_tensor(a)
                       else torch.as_tensor(np.ascontiguousarray(a, np.float32), device=self.dev))
        # projector P_D (make_D_projector, CDL_setup.jl:15-27) and the masked filters (:183)
        P = np.zeros((M, Lg), np.float32)
        for m, f in enumerate(self.fl):
            P[m, :f] = 1.0
        self.P = T(P)
        D = self.P * T(D0)
        # X = Pi4(randn) (:191-192), here Pi4(X0)
        X = T(X0).clone()
        keep = torch.zeros(Lg, dtype=f32, device=self.dev)
        keep[0::4] = 1.0
        self.keep4 = keep
        self.X = X * keep
        Dh = torch.fft.fft(D, dim=-1)
        self.Xh = torch.fft.fft(self.X, dim=-1)
        self.DX = torch.fft.ifft((Dh.unsqueeze(0) * self.Xh).sum(dim=1), dim=-1).real      # :201-203
        self.dh = Dh.unsqueeze(0).expand(N, M, Lg).clone()                                   # :207-208
        self.Td = torch.zeros((N, Lg), dtype=f32, device=self.dev)
        self.d = torch.fft.ifft(self.dh, dim=-1).real.contiguous()
        # W'W diagonal and W's (:213-217)
        self.WtW = torch.zeros(Lg, dtype=f32, device=self.dev)
        self.WtW[g:g + self.L4] = 1.0
        self.WtS = torch.zeros((N, Lg), dtype=f32, device=self.dev)
        self.WtS[:, g:g + self.L4] = T(s)
        self.s = T(s)
        # CSC auxiliaries and multipliers (:220-224)
        self.Y1 = torch.zeros((N, Lg), dtype=f32, device=self.dev)
        self.Y2 = self.X.clone()
        self.Y3 = self.X.clone()
        self.Y1h = torch.fft.fft(self.Y1, dim=-1)
        self.Y2h = torch.fft.fft(self.Y2, dim=-1)
        self.Y3h = self.Y2h.clone()
        self.y1 = torch.zeros_like(self.Y1)
        self.y2 = torch.zeros_like(self.X)
        self.y3 = torch.zeros_like(self.X)
        self.y1h = self.Y1h.clone()
        self.y2h = torch.fft.fft(self.y2, dim=-1)
        self.y3h = self.y2h.clone()
        # dictionary auxiliaries and multipliers (:226-230)
        self.G1 = D.clone()
        self.G2 = self.DX.clone()
        self.G3 = D.clone()
        self.G1h = torch.fft.fft(self.G1, dim=-1)
        self.G2h = torch.fft.fft(self.G2, dim=-1)
        self.G3h = self.G1h.clone()
        self.g1 = torch.zeros_like(self.X)
        self.g2 = torch.zeros_like(self.Y1)
        self.g3 = torch.zeros_like(self.X)
        self.g1h = self.y2h.clone()
        self.g2h = self.Y1h.clone()
        self.g3h = self.y2h.clone()
        self.Dbest = self.G1.clone()
        self.Xbest = self.Y3.clone()
        self.obj_best = float(np.float32(self.objective()))                  # :236, a Float32 field
        self.sync()

    def sync(self):
        if self.dev.type == "cuda":
            self.t.cuda.synchronize(self.dev)

    def objective(self) -> float:
        """0.5 norm(W DX - s)^2 + lambda sum|X| (CDL_setup.jl:286): Float32 norm and sum,
        Float64 0.5 and final addition."""
        t = self.t
        g = self.gamma
        r = self.DX[:, g:g + self.L4] - self.s
        nrm = float(t.linalg.vector_norm(r))
        nrm = float(np.float32(nrm))
        return 0.5 * float(np.float32(nrm) * np.float32(nrm)) + float(np.float32(self.lam) * np.float32(float(self.X.abs().sum())))

    def code_update(self, num_iter: int, sherman: float = 2.0):
        """code_update! (train.jl:1-44)."""
        t = self.t
        fft, ifft = t.fft.fft, t.fft.ifft
        for _ in range(num_iter):
            tmp = self.X + self.y2                                              # shrink! (:7)
            self.Y2 = t.sign(tmp) * t.clamp(tmp.abs() - self.lam, min=0.0)
            self.sync()
            self.Y2h = fft(self.Y2, dim=-1)                                     # :8
            self.y1 = self.y1 + self.DX - self.Y1                               # :12
            self.y1h = fft(self.y1, dim=-1)
            self.y2 = self.y2 + self.X - self.Y2                                # :13
            self.y2h = fft(self.y2, dim=-1)
            self.y3 = self.y3 + self.X - self.Y3                                # :14
            self.y3h = fft(self.y3, dim=-1)
            dots = self.G1h * self.G1h.conj()                                   # :18-20
            sum_dots = dots.sum(dim=0, keepdim=True) + sherman
            A = self.G1h / sum_dots
            RHS = (self.G1h.conj().unsqueeze(0) * (self.Y1h - self.y1h).unsqueeze(1)      # sherman_CSC_RHS! (:23)
                   + (self.Y2h - self.y2h) + (self.Y3h - self.y3h))
            self.sync()
            alpha = (A.unsqueeze(0) * RHS).sum(dim=1)                           # sherman_CSC_dot! (:24)
            self.sync()
            self.Xh = (RHS - alpha.unsqueeze(1) * self.G1h.conj().unsqueeze(0)) / sherman   # combine (:25)
            self.sync()
            self.X = ifft(self.Xh, dim=-1).real.contiguous()                    # :26
            self.DX = ifft((self.G1h.unsqueeze(0) * self.Xh).sum(dim=1), dim=-1).real      # :27
            self.Y1 = (self.WtS + self.rho * (self.DX + self.y1)) / (self.WtW + self.rho)  # update_Y1! (:35)
            self.sync()
            self.Y1h = fft(self.Y1, dim=-1)                                     # :36
            self.Y3 = (self.X + self.y3) * self.keep4                           # projectionX! (:40)
            self.sync()
            self.Y3h = fft(self.Y3, dim=-1)                                     # :41

    def dict_update(self, num_iter: int, sherman: float = 2.0):
        """dict_update! (train.jl:46-89)."""
        t = self.t
        fft, ifft = t.fft.fft, t.fft.ifft
        N = self.N
        for _ in range(num_iter):
            dots = self.Y3h * self.Y3h.conj()                                   # :52-54
            sum_dots = dots.sum(dim=1, keepdim=True) + sherman
            Xc = self.Y3h / sum_dots
            self.Td = ifft((self.Y3h * self.dh).sum(dim=1), dim=-1).real        # :56-57
            RHS = (self.Y3h.conj() * (self.G2h - self.g2h).unsqueeze(1)         # sherman_CDL_RHS! (:58)
                   + (self.G1h.unsqueeze(0) - self.g1h) + (self.G3h.unsqueeze(0) - self.g3h))
            self.sync()
            beta = (Xc * RHS).sum(dim=1)                                         # :59
            self.dh = RHS - beta.unsqueeze(1) * self.Y3h.conj()                  # sherman_CDL_combine (:60)
            self.sync()
            self.d = ifft(self.dh, dim=-1).real.contiguous()                     # :61
            self.G2 = (self.WtS + self.sigma * (self.Td + self.g2)) / (self.WtW + self.sigma)   # update_g2! (:68)
            self.sync()
            self.G2h = fft(self.G2, dim=-1)
            self.g1 = self.g1 + self.d - self.G1.unsqueeze(0)                    # :70-72
            self.g1h = fft(self.g1, dim=-1)
            self.g2 = self.g2 + self.Td - self.G2
            self.g2h = fft(self.g2, dim=-1)
            self.g3 = self.g3 + self.d - self.G3.unsqueeze(0)
            self.g3h = fft(self.g3, dim=-1)
            v2 = (self.d + self.g3).sum(dim=0) / N                               # :76
            G3 = t.where(v2 >= 1, (self.P != 0).to(v2.dtype), t.where(v2 <= 0, t.zeros_like(v2), v2 * (self.P != 0)))
            self.G3 = G3                                                         # projection2box! (:77)
            self.sync()
            self.G3h = fft(self.G3, dim=-1)
            v1 = (self.d + self.g1).sum(dim=0) / N                               # :83
            # prox_simplex_zeros! (cdl_kernels.jl:53-66): group i = 4 floor(l/4) of the 1-based l,
            # the 0.25 (...) term in Float64
            Lg = self.Lg
            l1 = t.arange(1, Lg + 1, device=self.dev)
            i0 = 4 * t.div(l1, 4, rounding_mode="floor")                        # 1-based i; v[i+1..i+4]
            idx = (i0.unsqueeze(0) + t.arange(0, 4, device=self.dev).unsqueeze(1)).clamp(max=Lg - 1)   # 0-based i..i+3
            grp = v1[:, idx].sum(dim=1) - 1.0                                    # (M, Lg), Float32
            G1 = (v1.double() - 0.25 * grp.double()).float()
            self.G1 = t.where(self.P != 0, G1, t.zeros_like(G1))
            self.sync()
            self.G1h = fft(self.G1, dim=-1)                                      # :85

    def update_best(self):
        """update_best! (train.jl:91-98); returns (objective, improved)."""
        obj = self.objective()
        imp = obj < self.obj_best
        if imp:
            self.obj_best = float(np.float32(obj))
            self.Dbest = self.d.sum(dim=0) / self.N
            self.Xbest = self.X.clone()
        return obj, imp

    def get(self, name: str) -> np.ndarray:
        return getattr(self, name).detach().cpu().numpy()
