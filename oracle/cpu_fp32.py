"""CPU BASELINE — the reference's exhaustive ccc-only template-matching loop in
FP32 on the host's threads (test infrastructure: only bench.py's
cpu_baseline leg and tests/ use it; the product never imports it).

This is the loop of Oscar3/matlabPicker/search/tom_os3_findTemplate.m:157-307
with Oscar3/matlabPicker/search/tom_os3_corr.m:77-192 (FLCF branch) as the
reference runs it, in single precision (tom_os3_corr.m:56-57 casts to
`single`), with PyTorch's CPU FFTs (MKL / pocketfft) on
torch.get_num_threads() threads:

  once:   F_I = fftn(I); mu_I, sigma_I by FFT convolution with the pasted mask
          (statistics/tom_os3_mean.m:44-63, tom_os3_std.m:43-68), FP32;
          template tom_norm(1e5*(T+1), 'mean0+1std') and its fixed statistics
          (tom_os3_findTemplate.m:157, 233-234, tom_os3_normUnderMask.m)
  per orientation a:
          t = tom_rotate(T, a) .* M        (rot3d float arithmetic, C)
          t = tom_apply_weight_function(t, W) .* M   (box size, FP64 as the
                                                       reference's :56 does)
          c = real(ifftshift(ifftn(F_I .* conj(fftn(pasteCenter(t))))))
          score = (c/m - mu_I mu_T) / (sigma_T sigma_I)   (0 where sigma_I = 0)
          bestPeak update: strict '>' keeps the first angle (tom_os3_bestPeak.m:94-98)

The pasted template sits with its centre voxel at the origin (wrap-around),
which equals pasteCenter followed by ifftshift for even sizes. Pinned against
the FP64 oracle in tests/test_oracle.py (test_cpu_fp32_baseline_matches_oracle).
"""
from __future__ import annotations

import time

import numpy as np


class CpuFlcfFP32:
    """FP32 exhaustive FLCF on the CPU for one volume, template, mask, weight."""

    def __init__(self, V, T, M, W=None):
        import torch

        from oracle import tom_oracle as O
        self.torch = torch
        self.O = O
        V = np.asarray(V, np.float32)
        self.N = V.shape
        M = np.asarray(M, np.float32)
        self.M = M
        self.W = None if W is None else np.asarray(W, np.float64)
        self.m = float((M > 0).sum())
        n = M.shape
        self.n = n
        # template normalisation and its fixed statistics (FP64 at box size,
        # one-off; tom_os3_findTemplate.m:157, 233-234)
        self.Tn, self.mu_T, self.sig_T = O.prepare_template(T, M.astype(np.float64), self.W)
        self.Tn32 = np.asarray(self.Tn, np.float32)
        t0 = time.perf_counter()
        I = torch.from_numpy(np.ascontiguousarray(V))
        self.FI = torch.fft.rfftn(I)
        Mp = torch.zeros(self.N, dtype=torch.float32)
        self._place(Mp, torch.from_numpy(M))
        FM = torch.fft.rfftn(Mp)
        mu = torch.fft.irfftn(self.FI * FM, s=self.N) / self.m
        e2 = torch.fft.irfftn(torch.fft.rfftn(I * I) * FM, s=self.N) / self.m
        var = e2 - mu * mu
        self.sig_I = torch.sqrt(torch.clamp(var, min=0.0))
        self.mu_I = mu
        self.precompute_s = time.perf_counter() - t0
        self.best = torch.full(self.N, -1000.0, dtype=torch.float32)
        self.arg = torch.zeros(self.N, dtype=torch.int32)
        self.count = 0

    def _place(self, dst, t):
        """Template (n1, n2, n3) into dst with its centre voxel floor(n/2) at
        the origin (wrap-around): pasteCenter + ifftshift."""
        n = t.shape
        c = [k // 2 for k in n]
        idx = [(np.arange(k) - ck) % Nk for k, ck, Nk in zip(n, c, dst.shape)]
        dst[np.ix_(*idx)] = t

    def template(self, angle):
        """Rotated, masked, weighted, re-masked template (tom_os3_findTemplate.m:264-266)."""
        t = self.O.rotate(self.Tn32, angle).astype(np.float64) * self.M
        if self.W is not None:
            t = self.O.apply_weight(t, self.W)
        return (t * self.M).astype(np.float32)

    def score(self, angle):
        torch = self.torch
        tp = torch.zeros(self.N, dtype=torch.float32)
        self._place(tp, torch.from_numpy(self.template(angle)))
        c = torch.fft.irfftn(self.FI * torch.conj(torch.fft.rfftn(tp)), s=self.N)
        den = self.sig_I * float(self.sig_T)
        num = c / self.m - self.mu_I * float(self.mu_T)
        return torch.where(den > 0, num / torch.where(den > 0, den, torch.ones_like(den)), torch.zeros_like(den))

    def step(self, angle):
        """One orientation of the search loop with the bestPeak update
        (1-based angle index, maps start at -1000)."""
        torch = self.torch
        s = self.score(angle)
        upd = self.best < s
        self.best = torch.where(upd, s, self.best)
        self.arg = torch.where(upd, torch.full_like(self.arg, self.count + 1), self.arg)
        self.count += 1


def cpu_info():
    """(logical CPUs, model name) of this host."""
    import os
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count(), model
