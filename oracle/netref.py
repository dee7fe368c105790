"""FP64 numpy restatement of the learned-init network (test infrastructure).

The network has no counterpart in the reference; this module is the checker
for the GPU kernels in poisson_caustic_design_amd/csrc/cd_learned.hip.  It
builds the same layer list in the same order, uses the same flat parameter
layout ([coutpad][kpad] per convolution, k = tap * Cin + ci, tap = 3*dy + dx)
and computes forward, loss and gradients in float64.  `emulate_bf16` rounds
activations and weights to BF16 the way the GPU stores them, for tight
forward comparisons.  Its gradients are pinned by central finite differences,
its layers and their wiring by closed forms (impulse responses, SciPy's
cross-correlation, adjoint pairs, identity weights that reduce the network to
block means or to the identity) in tests/test_netref.py.
"""
from __future__ import annotations

import numpy as np


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


class Conv:
    def __init__(self, c1, c2, cout, cout_alloc, relu, level, in1, in2, out, woff):
        self.c1, self.c2, self.cout, self.coutA, self.relu = c1, c2, cout, cout_alloc, relu
        self.level, self.in1, self.in2, self.out = level, in1, in2, out
        self.K = 9 * (c1 + c2)
        self.kpad = (self.K + 31) // 32 * 32
        self.coutpad = (cout_alloc + 15) // 16 * 16
        self.woff = woff
        self.nparam = self.coutpad * self.kpad


class NetRef:
    def __init__(self, levels=4, channels=(16, 32, 64, 64), linear=False, out_smooth=0,
                 loss_smooth=0, loss_energy=False):
        self.L = levels
        self.C = list(channels)[:levels]
        self.act = 0 if linear else 1
        self.out_smooth = out_smooth    # y = S^k head, S = I + Lc/8
        self.loss_smooth = loss_smooth  # operator loss on S^k (cs Lc y - x)
        self.loss_energy = loss_energy  # energy y'(-cs Lc)y/2 + x'y instead
        self.convs: list[Conv] = []
        self.tensors: list[tuple[int, int]] = []  # (level, channels)
        self.nparam = 0
        self._build()

    def _tensor(self, level, ch):
        self.tensors.append((level, ch))
        return len(self.tensors) - 1

    def _conv(self, in1, c1, in2, c2, out, cout, coutA, relu, level):
        cv = Conv(c1, c2, cout, coutA, relu, level, in1, in2, out, self.nparam)
        self.nparam += cv.nparam
        self.convs.append(cv)

    def _build(self):
        L, C = self.L, self.C
        skip = [0] * L
        prev, prevC = -1, 8
        for lv in range(L):
            inp, inC = prev, prevC
            if lv > 0:
                inp = self._tensor(lv, prevC)
            a = self._tensor(lv, C[lv])
            self._conv(inp, inC, -1, 0, a, C[lv], C[lv], self.act, lv)
            s = self._tensor(lv, C[lv])
            self._conv(a, C[lv], -1, 0, s, C[lv], C[lv], self.act, lv)
            skip[lv] = s
            prev, prevC = s, C[lv]
        d, dC = skip[L - 1], C[L - 1]
        for lv in range(L - 2, -1, -1):
            u = self._tensor(lv, dC)
            b = self._tensor(lv, C[lv])
            self._conv(u, dC, skip[lv], C[lv], b, C[lv], C[lv], self.act, lv)
            e = self._tensor(lv, C[lv])
            self._conv(b, C[lv], -1, 0, e, C[lv], C[lv], self.act, lv)
            d, dC = e, C[lv]
        self._conv(d, dC, -1, 0, -1, 1, 8, 0, 0)

    # ---- parameters ----------------------------------------------------
    def weights(self, params, cv: Conv, emulate_bf16=False):
        """[9][Cin][coutA] tap-major weight tensor of one convolution."""
        w = params[cv.woff:cv.woff + cv.nparam].reshape(cv.coutpad, cv.kpad)
        w = w[:cv.coutA, :cv.K].astype(np.float64)
        if emulate_bf16:
            w = bf16(w)
        cin = cv.c1 + cv.c2
        return w.reshape(cv.coutA, 9, cin).transpose(1, 2, 0)

    def init_params(self, seed=0):
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparam, dtype=np.float32)
        for i, cv in enumerate(self.convs):
            cin = cv.c1 + cv.c2
            fan = 9 * (1 if i == 0 else cin)
            if self.act:
                std = np.sqrt(2.0 / fan) * (0.1 if cv.out < 0 else 1.0)
            else:
                std = np.sqrt(1.0 / (3.0 * fan)) * (0.0 if cv.out < 0 else 1.0)
            w = np.zeros((cv.coutpad, cv.kpad), dtype=np.float32)
            vals = rng.standard_normal((cv.cout, cv.K)).astype(np.float32) * std
            if i == 0:
                vals[:, [k for k in range(cv.K) if k % cin != 0]] = 0.0
            w[:cv.cout, :cv.K] = vals
            p[cv.woff:cv.woff + cv.nparam] = w.ravel()
        return p

    def random_params(self, seed=0, head_scale=0.3):
        """Random weights in every layer (head included), for gradient and
        parity checks that must not start from a zero head."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparam, dtype=np.float32)
        for cv in self.convs:
            cin = cv.c1 + cv.c2
            std = np.sqrt(1.0 / (9 * cin)) * (head_scale if cv.out < 0 else 1.0)
            w = np.zeros((cv.coutpad, cv.kpad), dtype=np.float32)
            w[:cv.cout, :cv.K] = rng.standard_normal((cv.cout, cv.K)).astype(np.float32) * std
            p[cv.woff:cv.woff + cv.nparam] = w.ravel()
        return p

    # ---- primitives ----------------------------------------------------
    @staticmethod
    def conv(x, wt):
        """x [B,H,W,Cin], wt [9][Cin][Cout] -> [B,H,W,Cout], zero padding."""
        B, H, W, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        y = np.zeros((B, H, W, wt.shape[2]))
        for tap in range(9):
            dy, dx = tap // 3, tap % 3
            y += xp[:, dy:dy + H, dx:dx + W, :] @ wt[tap]
        return y

    @staticmethod
    def conv_bwd(x, wt, g):
        """gradients of conv w.r.t. x and wt for upstream g [B,H,W,Cout]."""
        B, H, W, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        dxp = np.zeros_like(xp)
        dw = np.zeros_like(wt)
        g2 = g.reshape(-1, g.shape[-1])
        for tap in range(9):
            dy, dx = tap // 3, tap % 3
            xs = xp[:, dy:dy + H, dx:dx + W, :]
            dw[tap] = xs.reshape(-1, xs.shape[-1]).T @ g2
            dxp[:, dy:dy + H, dx:dx + W, :] += g @ wt[tap].T
        return dxp[:, 1:-1, 1:-1, :], dw

    @staticmethod
    def pool(x):
        B, H, W, C = x.shape
        return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))

    @staticmethod
    def pool_bwd(g):
        return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) / 4.0

    @staticmethod
    def up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    @staticmethod
    def up_bwd(g):
        B, H, W, C = g.shape
        return g.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))

    # ---- network -------------------------------------------------------
    def forward(self, params, x, emulate_bf16=False):
        """x [B,H,W] -> (y [B,H,W], cache)."""
        rd = bf16 if emulate_bf16 else (lambda a: a)
        B, H, W = x.shape
        x8 = np.zeros((B, H, W, 8))
        x8[..., 0] = rd(x)
        acts = [None] * len(self.tensors)
        L = self.L
        cache = []
        y = None
        for idx, cv in enumerate(self.convs):
            xin = x8 if cv.in1 < 0 else acts[cv.in1]
            if cv.in2 >= 0:
                xin = np.concatenate([xin, acts[cv.in2]], axis=-1)
            wt = self.weights(params, cv, emulate_bf16)
            pre = self.conv(xin, wt)
            out = np.maximum(pre, 0.0) if cv.relu else pre
            cache.append((xin, wt, out))
            if cv.out < 0:
                y = self.smooth(out[..., 0], self.out_smooth)
            else:
                acts[cv.out] = rd(out)
                if idx < 2 * L and idx % 2 == 1 and cv.level < L - 1:
                    acts[self.convs[idx + 1].in1] = rd(self.pool(acts[cv.out]))
                if idx + 1 < len(self.convs) and self.convs[idx + 1].in2 >= 0:
                    acts[self.convs[idx + 1].in1] = self.up(acts[cv.out])
        return y, (acts, cache)

    @classmethod
    def smooth(cls, v, k):
        """k steps of v += Lc v / 8 (symmetric, gain 1 - lambda/8)."""
        for _ in range(k):
            v = v + cls.lap(v) / 8.0
        return v

    def loss_grad(self, params, x, t, emulate_bf16=False):
        """Mean squared error against targets t and its parameter gradient."""
        y, cache = self.forward(params, x, emulate_bf16)
        loss = float(np.mean((y - t) ** 2))
        return loss, self._backward(cache, 2.0 * (y - t) / y.size)

    @staticmethod
    def lap(v):
        """Neumann 5-point operator (sum of in-grid neighbours - count * centre)
        on the last two axes."""
        out = np.zeros_like(v)
        cnt = np.zeros(v.shape[-2:])
        out[..., :, 1:] += v[..., :, :-1]; cnt[:, 1:] += 1
        out[..., :, :-1] += v[..., :, 1:]; cnt[:, :-1] += 1
        out[..., 1:, :] += v[..., :-1, :]; cnt[1:, :] += 1
        out[..., :-1, :] += v[..., 1:, :]; cnt[:-1, :] += 1
        return out - cnt * v

    def loss_grad_operator(self, params, x, emulate_bf16=False):
        """Operator loss mean((S^k (cs * Lc y - x))^2), cs = (max(H, W)/pi)^2,
        (or, with loss_energy, the energy mean(y (x - u) / 2), u = cs Lc y - x),
        k = loss_smooth, and its parameter gradient (cs Lc S^k (2 v / n) into
        y; Lc and S are symmetric)."""
        y, cache = self.forward(params, x, emulate_bf16)
        cs = (max(x.shape[1], x.shape[2]) / np.pi) ** 2
        if self.loss_energy:
            # J = mean(y (x - u) / 2), u = cs Lc y - x; dJ/dy = -u / n
            u = cs * self.lap(y) - x
            return float(np.mean(0.5 * y * (x - u))), self._backward(cache, -u / y.size)
        v = self.smooth(cs * self.lap(y) - x, self.loss_smooth)
        loss = float(np.mean(v ** 2))
        dv = 2.0 * v / y.size
        return loss, self._backward(cache, cs * self.lap(self.smooth(dv, self.loss_smooth)))

    def _backward(self, fwd_cache, g_y):
        acts, cache = fwd_cache
        g_t = [np.zeros_like(a) if a is not None else None for a in acts]
        gflat = np.zeros(self.nparam)
        L = self.L
        for idx in range(len(self.convs) - 1, -1, -1):
            cv = self.convs[idx]
            xin, wt, out = cache[idx]
            if cv.out < 0:
                g = np.zeros(out.shape)
                g[..., 0] = self.smooth(g_y, self.out_smooth)  # S is symmetric
            else:
                g = g_t[cv.out] * (out > 0 if cv.relu else 1.0)
            dx, dw = self.conv_bwd(xin, wt, g)
            # scatter dw [9][Cin][coutA] into the flat [coutpad][kpad] layout
            cin = cv.c1 + cv.c2
            blk = np.zeros((cv.coutpad, cv.kpad))
            blk[:cv.coutA, :cv.K] = dw.transpose(2, 0, 1).reshape(cv.coutA, 9 * cin)
            gflat[cv.woff:cv.woff + cv.nparam] = blk.ravel()
            if cv.in1 >= 0:
                g_t[cv.in1] += dx[..., :cv.c1]
                if cv.in2 >= 0:
                    g_t[cv.in2] += dx[..., cv.c1:]
                    pv = self.convs[idx - 1]
                    g_t[pv.out] += self.up_bwd(g_t[cv.in1])
                if idx < 2 * L and idx % 2 == 0 and cv.level > 0:
                    pv = self.convs[idx - 1]
                    g_t[pv.out] += self.pool_bwd(g_t[cv.in1])
        return gflat
