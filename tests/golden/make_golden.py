"""Writes the golden fixtures under tests/golden/ from an independent numpy restatement.

The reference (CDLmotif.jl, pure Julia + CUDA.jl) cannot run here and ships no test
vectors (test/runtests.jl:4-6 is an empty testset).  This script is a second,
independently written restatement of the same algorithm -- vectorised numpy in float64
with numpy's pocketfft -- so that agreement between it and the C oracle (loop code with
its own FFT) pins the oracle's transcription.  Sources restated:
  src/SETUP/CDL_setup.jl:15-40,165-236,286   init and objective
  src/CDL/train.jl:1-98                      code_update!, dict_update!, update_best!
  src/CDL/csc_kernels.jl:8-95, cdl_kernels.jl:2-77
  src/WRAPPER/wrap.jl:4-12                   schedule and code threshold
  src/SEARCH/scan.jl:10-21,89-121,176-239    PWM scans
Run:  python tests/golden/make_golden.py   (writes cdl_*.npz and pwm_*.npz here)
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------------------- CDL (float64)
class NpCDL:
    def __init__(self, s, D0, X0, fl, rho, sigma, lam):
        s = s.astype(np.float64)                      # (N, L4)
        self.N, self.L4 = s.shape
        self.fl = np.asarray(fl)
        self.M = len(fl)
        self.g = int(self.fl.max())
        self.Lg = self.L4 + 2 * self.g
        self.rho, self.sigma, self.lam = rho, sigma, lam
        Lg = self.Lg
        l = np.arange(Lg)
        self.P = (l[None, :] < self.fl[:, None]).astype(np.float64)      # (M, Lg)
        D = self.P * D0.astype(np.float64)
        X = X0.astype(np.float64) * (l % 4 == 0)                          # (N, M, Lg)
        F = lambda a: np.fft.fft(a, axis=-1)
        iF = lambda a: np.fft.ifft(a, axis=-1).real
        self.F, self.iF = F, iF
        self.s = s
        self.w = ((l >= self.g) & (l < self.g + self.L4)).astype(np.float64)
        self.b = np.zeros((self.N, Lg))
        self.b[:, self.g:self.g + self.L4] = s
        Dh = F(D)
        self.X = X
        self.DX = iF((Dh[None] * F(X)).sum(axis=1))
        self.dh = np.broadcast_to(Dh, (self.N, self.M, Lg)).copy()
        self.d = iF(self.dh)
        self.Td = np.zeros((self.N, Lg))
        self.Y1 = np.zeros((self.N, Lg)); self.y1 = np.zeros((self.N, Lg))
        self.Y3 = X.copy(); self.y2 = np.zeros_like(X); self.y3 = np.zeros_like(X)
        self.Y1h = F(self.Y1); self.Y3h = F(self.Y3)
        self.G1 = D.copy(); self.G3 = D.copy(); self.G2 = self.DX.copy()
        self.g1 = np.zeros_like(X); self.g3 = np.zeros_like(X); self.g2 = np.zeros((self.N, Lg))
        self.G1h, self.G3h, self.G2h = F(self.G1), F(self.G3), F(self.G2)
        self.g1h = np.zeros(X.shape, complex); self.g3h = np.zeros(X.shape, complex)
        self.g2h = np.zeros((self.N, Lg), complex)
        self.Xbest = X.copy(); self.Dbest = D.copy()
        self.obj_best = self.objective()

    def objective(self):
        r = self.DX[:, self.g:self.g + self.L4] - self.s
        return 0.5 * np.sum(r * r) + self.lam * np.abs(self.X).sum()

    def code_update(self, iters, sh=2.0):
        F, iF = self.F, self.iF
        for _ in range(iters):
            t = self.X + self.y2
            Y2 = np.sign(t) * np.maximum(np.abs(t) - self.lam, 0)
            self.y1 = self.y1 + self.DX - self.Y1
            self.y2 = self.y2 + self.X - Y2
            self.y3 = self.y3 + self.X - self.Y3
            G = self.G1h[None]                                               # (1, M, Lg)
            c = (np.abs(self.G1h) ** 2).sum(axis=0) + sh                       # (Lg,)
            R = np.conj(G) * F(self.Y1 - self.y1)[:, None, :] + F(Y2 - self.y2) + F(self.Y3 - self.y3)
            alpha = ((G / c) * R).sum(axis=1)                                   # (N, Lg)
            Xh = (R - alpha[:, None, :] * np.conj(G)) / sh
            self.X = iF(Xh)
            self.DX = iF((G * Xh).sum(axis=1))
            self.Y1 = (self.b + self.rho * (self.DX + self.y1)) / (self.w + self.rho)
            l = np.arange(self.Lg)
            self.Y3 = (self.X + self.y3) * (l % 4 == 0)

    def dict_update(self, iters, sh=2.0):
        F, iF = self.F, self.iF
        for _ in range(iters):
            Y3h = F(self.Y3)                                                     # (N, M, Lg)
            c = (np.abs(Y3h) ** 2).sum(axis=1) + sh                              # (N, Lg)
            self.Td = iF((Y3h * self.dh).sum(axis=1))
            R = np.conj(Y3h) * (self.G2h - self.g2h)[:, None, :] + (self.G1h[None] - self.g1h) + (self.G3h[None] - self.g3h)
            beta = ((Y3h / c[:, None, :]) * R).sum(axis=1)
            self.dh = R - beta[:, None, :] * np.conj(Y3h)
            self.d = iF(self.dh)
            G2 = (self.b + self.sigma * (self.Td + self.g2)) / (self.w + self.sigma)
            self.g1 = self.g1 + self.d - self.G1[None]
            self.g2 = self.g2 + self.Td - G2
            self.g3 = self.g3 + self.d - self.G3[None]
            self.G2 = G2
            self.G2h, self.g1h, self.g2h, self.g3h = F(G2), F(self.g1), F(self.g2), F(self.g3)
            v2 = (self.d + self.g3).sum(axis=0) / self.N
            self.G3 = np.where(v2 >= 1, 1.0, np.where(v2 <= 0, 0.0, v2)) * (self.P != 0)
            self.G3h = F(self.G3)
            v1 = (self.d + self.g1).sum(axis=0) / self.N
            # group start 4*floor(l/4) for 1-based l  -> 0-based rows i..i+3, i = 4*((l0+1)//4)
            l0 = np.arange(self.Lg)
            i = np.minimum(4 * ((l0 + 1) // 4), self.Lg - 4)   # clipping only touches off-support rows
            grp = v1[:, i] + v1[:, i + 1] + v1[:, i + 2] + v1[:, i + 3]
            self.G1 = np.where(self.P != 0, v1 - 0.25 * (grp - 1), 0.0)
            self.G1h = F(self.G1)

    def update_best(self):
        o = self.objective()
        if o < self.obj_best:
            self.obj_best = o
            self.Dbest = self.d.sum(axis=0) / self.N
            self.Xbest = self.X.copy()
            return o, True
        return o, False

    def code_thresh(self, mult=1.0):
        v = self.Xbest[self.Xbest > 0]
        return mult * np.std(v, ddof=1), v.size


def make_cdl_inputs(seed, N, L, M, fl, n_motif=1):
    """One-hot data (N, 4L) with planted motifs; D0 (M, Lg) masked rand_probvec groups
    (CDL_setup.jl:59-68,183); X0 (N, M, Lg) standard normal."""
    rng = np.random.default_rng(seed)
    bases = rng.integers(0, 4, (N, L))
    for _ in range(n_motif):
        m = min(6, L)
        pfm = rng.dirichlet(0.5 * np.ones(4), m)
        for n in range(N):
            if rng.random() < 0.5:
                at = rng.integers(0, L - m + 1)
                bases[n, at:at + m] = [rng.choice(4, p=p) for p in pfm]
    s = np.zeros((N, 4 * L), np.float32)
    s[np.arange(N)[:, None], 4 * np.arange(L)[None, :] + bases] = 1
    g = max(fl)
    Lg = 4 * L + 2 * g
    D0 = np.zeros((M, Lg), np.float32)
    for m in range(M):
        u = np.sort(rng.random((Lg // 4, 3)), axis=1)
        D0[m] = np.diff(np.concatenate([np.zeros((Lg // 4, 1)), u, np.ones((Lg // 4, 1))], axis=1), axis=1).ravel()
    X0 = rng.standard_normal((N, M, Lg)).astype(np.float32)
    return s, D0, X0, bases.astype(np.uint8)


CDL_CASES = [
    # name, seed, N, L, M, fl, rho, sigma, lam, schedule (outer, csc, dict)
    ("cdl_tiny", 11, 3, 2, 2, [8, 8], 1.0, 100.0, 0.05, (2, 3, 1)),
    ("cdl_small", 12, 5, 10, 3, [8, 8, 8], 1.0, 100.0, 0.05, (4, 5, 1)),
    ("cdl_mixedfl", 13, 4, 8, 3, [4, 8, 12], 100.0, 100.0, 0.01, (3, 4, 1)),
    ("cdl_long", 14, 3, 6, 2, [8, 8], 1.0, 100.0, 0.05, (10, 15, 1)),
]


def run_case(name, seed, N, L, M, fl, rho, sigma, lam, sched):
    s, D0, X0, bases = make_cdl_inputs(seed, N, L, M, fl)
    c = NpCDL(s, D0, X0, fl, rho, sigma, lam)
    outer, ncsc, ndict = sched
    objs, imps = [], []
    for _ in range(outer):
        c.code_update(ncsc)
        o, imp = c.update_best()
        objs.append(o); imps.append(imp)
        c.dict_update(ndict)
    th, cnt = c.code_thresh(1.0)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), s=s, D0=D0, X0=X0, fl=np.asarray(fl, np.int32),
                        params=np.array([rho, sigma, lam]), sched=np.array(sched), X=c.X, DX=c.DX, d=c.d, G1=c.G1,
                        G3=c.G3, Xbest=c.Xbest, Dbest=c.Dbest, y1=c.y1, y2=c.y2, y3=c.y3, Y1=c.Y1, g1=c.g1,
                        g2=c.g2, g3=c.g3, Td=c.Td, obj=np.array(objs), improved=np.array(imps), obj_best=c.obj_best,
                        thresh=th, thresh_count=cnt)


# ----------------------------------------------------------------------------- PWM
def pwm_cases():
    rng = np.random.default_rng(77)
    N, L, K = 40, 60, 7
    lens = np.array([1, 3, 6, 8, 11, 15, 24], np.int32)
    maxlen = int(lens.max())
    pwms = np.zeros((K, 4, maxlen), np.float32)
    for k in range(K):
        p = rng.dirichlet(0.5 * np.ones(4), lens[k])
        p = (p + 0.01) / 1.04
        pwms[k, :, :lens[k]] = np.log2(p / 0.25).T.astype(np.float32)
    bases = rng.choice(4, size=(N, L), p=[0.29, 0.21, 0.21, 0.29]).astype(np.uint8)
    # plant a few perfect consensus copies on both strands
    for k in range(K):
        cons = np.argmax(pwms[k, :, :lens[k]], axis=0)
        for n in rng.choice(N, 3, replace=False):
            at = rng.integers(0, L - lens[k] + 1)
            bases[n, at:at + lens[k]] = cons if rng.random() < 0.5 else (3 - cons[::-1])
    best = np.array([np.float32(0)] * K)
    for k in range(K):
        acc = np.float32(0)
        for j in range(lens[k]):
            acc = np.float32(acc + pwms[k, :, j].max())
        best[k] = acc
    thresh = (np.float32(0.8) * best).astype(np.float32)
    thresh[0] = np.float32(-0.5)        # a negative threshold exercises the argmax quirk
    return bases, pwms, lens, thresh


def np_pwm_hits(bases, pwms, lens, thresh):
    out = []
    N, L = bases.shape
    for k in range(len(lens)):
        m = int(lens[k])
        fw = pwms[k, :, :m]
        rc = fw[::-1, ::-1]                       # pwm_complement = reverse in both dims
        for n in range(N):
            for l0 in range(L - m + 1):
                w = bases[n, l0:l0 + m]
                for st, P in ((0, fw), (1, rc)):
                    acc = np.float32(0)
                    for j in range(m):
                        acc = np.float32(acc + P[w[j], j])
                    if acc > thresh[k]:
                        out.append((k, n, l0, st, acc))
    return out


def np_pwm_gpu_argmax(bases, pwms, lens, thresh):
    N, L = bases.shape
    K = len(lens)
    pos = np.zeros((K, N), np.int32)
    for k in range(K):
        m = int(lens[k])
        for n in range(N):
            v = np.zeros(L, np.float32)
            for l0 in range(L - m + 1):
                acc = np.float32(0)
                for j in range(m):
                    acc = np.float32(acc + pwms[k, bases[n, l0 + j], j])
                v[l0] = acc if acc > thresh[k] else np.float32(0)
            a = int(np.argmax(v))
            pos[k, n] = 0 if a == 0 else a + 1
    return pos


def np_pwm_best(bases, pwms, lens, thresh):
    N, L = bases.shape
    K = len(lens)
    pos = np.zeros((K, N), np.int32); comp = np.zeros((K, N), np.int8); sc = np.zeros((K, N), np.float32)
    onehot = np.zeros((N, 4, L), np.float32)
    onehot[np.arange(N)[:, None], bases, np.arange(L)[None, :]] = 1
    for k in range(K):
        m = int(lens[k])
        fw = pwms[k, :, :m]
        rc = fw[::-1, ::-1]
        for n in range(N):
            res = []
            for P in (fw, rc):
                bi, bs = -1, -np.inf
                for i in range(L - m + 1):
                    score = np.float32(0)
                    for a in range(4):
                        for j in range(m):
                            score = np.float32(score + np.float32(P[a, j] * onehot[n, a, i + j]))
                    if score > bs:
                        bs, bi = score, i + 1
                res.append((bi, bs))
            (p, s), (pc, scc) = res
            if s >= scc:
                if s > thresh[k]:
                    pos[k, n], comp[k, n], sc[k, n] = p, 0, s
            elif scc > thresh[k]:
                pos[k, n], comp[k, n], sc[k, n] = pc, 1, scc
    return pos, comp, sc


def main():
    for case in CDL_CASES:
        run_case(*case)
        print("wrote", case[0])
    bases, pwms, lens, thresh = pwm_cases()
    hits = np_pwm_hits(bases, pwms, lens, thresh)
    h = np.array([(k, n, p, st) for k, n, p, st, _ in hits], np.int64).reshape(-1, 4)
    hs = np.array([s for *_, s in hits], np.float32)
    am = np_pwm_gpu_argmax(bases, pwms, lens, thresh)
    bp, bc, bs = np_pwm_best(bases, pwms, lens, thresh)
    np.savez_compressed(os.path.join(HERE, "pwm_small.npz"), bases=bases, pwms=pwms, lens=lens, thresh=thresh,
                        hits=h, hit_scores=hs, argmax_pos=am, best_pos=bp, best_comp=bc, best_score=bs)
    print("wrote pwm_small", len(hits), "hits")


if __name__ == "__main__":
    main()
