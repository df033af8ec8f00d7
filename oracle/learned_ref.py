"""FP64 numpy restatement of the learned-guess network (test infrastructure only).

Mirrors the parameter layout and arithmetic of the GPU kernels: NL conv
layers (kernel 5, stride 2, zero padding 2, ReLU), MLP flat -> H (ReLU) -> NB,
mean-squared-error loss over batch x NB.  Used to check the GPU gradients, and
checked itself against central finite differences.
"""
from __future__ import annotations

import numpy as np


class NetRef:
    def __init__(self, L0, C=32, NL=4, H=256, NB=64):
        self.L0, self.C, self.NL, self.H, self.NB = L0, C, NL, H, NB
        self.L = [L0]
        for _ in range(NL):
            self.L.append((self.L[-1] - 1) // 2 + 1)
        self.flat = C * self.L[NL]
        self.layout = []          # (name, shape) in parameter order
        for l in range(1, NL + 1):
            cin = 1 if l == 1 else C
            self.layout += [(f"w{l}", (C, cin, 5)), (f"b{l}", (C,))]
        self.layout += [("fc1w", (H, self.flat)), ("fc1b", (H,)), ("fc2w", (NB, H)), ("fc2b", (NB,))]
        self.nparams = sum(int(np.prod(s)) for _, s in self.layout)

    def unpack(self, p):
        out, o = {}, 0
        for name, shape in self.layout:
            n = int(np.prod(shape))
            out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
            o += n
        return out

    def pack(self, d):
        return np.concatenate([d[name].ravel() for name, _ in self.layout])

    @staticmethod
    def _windows(x):
        # x [B, C, L] -> [B, C, Lout, 5] with zero padding 2, stride 2
        B, Cc, L = x.shape
        Lout = (L - 1) // 2 + 1
        xp = np.zeros((B, Cc, L + 4))
        xp[:, :, 2:L + 2] = x
        idx = 2 * np.arange(Lout)[:, None] + np.arange(5)[None, :]
        return xp[:, :, idx]

    def forward(self, p, X):
        """Network outputs [B x NB] for inputs X [B x L0]."""
        P = self.unpack(p)
        a = np.asarray(X, dtype=np.float64)[:, None, :]
        for l in range(1, self.NL + 1):
            z = np.einsum("bcok,dck->bdo", self._windows(a), P[f"w{l}"]) + P[f"b{l}"][None, :, None]
            a = np.maximum(z, 0.0)
        h = np.maximum(a.reshape(a.shape[0], -1) @ P["fc1w"].T + P["fc1b"], 0.0)
        return h @ P["fc2w"].T + P["fc2b"]

    def loss_grad(self, p, X, Y):
        P = self.unpack(p)
        B = X.shape[0]
        acts = [np.asarray(X, dtype=np.float64)[:, None, :]]
        wins = []
        for l in range(1, self.NL + 1):
            xw = self._windows(acts[-1])
            wins.append(xw)
            z = np.einsum("bcok,dck->bdo", xw, P[f"w{l}"]) + P[f"b{l}"][None, :, None]
            acts.append(np.maximum(z, 0.0))
        flat = acts[-1].reshape(B, -1)
        h = np.maximum(flat @ P["fc1w"].T + P["fc1b"], 0.0)
        out = h @ P["fc2w"].T + P["fc2b"]
        diff = out - Y
        loss = float(np.mean(diff ** 2))
        g = {}
        dout = 2.0 * diff / diff.size
        g["fc2w"] = dout.T @ h
        g["fc2b"] = dout.sum(0)
        dh = (dout @ P["fc2w"]) * (h > 0)
        g["fc1w"] = dh.T @ flat
        g["fc1b"] = dh.sum(0)
        dflat = (dh @ P["fc1w"]) * (flat > 0)
        dy = dflat.reshape(acts[-1].shape)
        for l in range(self.NL, 0, -1):
            xw = wins[l - 1]
            g[f"w{l}"] = np.einsum("bdo,bcok->dck", dy, xw)
            g[f"b{l}"] = dy.sum((0, 2))
            if l > 1:
                dxw = np.einsum("bdo,dck->bcok", dy, P[f"w{l}"])
                xin = acts[l - 1]
                Lin = xin.shape[2]
                dxp = np.zeros((B, xin.shape[1], Lin + 4))
                Lout = dy.shape[2]
                for k in range(5):
                    np.add.at(dxp, (slice(None), slice(None), 2 * np.arange(Lout) + k), dxw[:, :, :, k])
                dy = dxp[:, :, 2:Lin + 2] * (xin > 0)
        return loss, self.pack(g)
