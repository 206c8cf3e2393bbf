"""Training data and training of the wave surrogate.

Data: the plain solver itself, run as a batched ensemble of seeded domains,
spun up, then averaged over wave-group windows of ``win_steps`` steps on the
device (xbg_ens_window_*).  Each window of each domain is one sample:
inputs zb and the window means of zs, uu, vv with the offshore Hm0, Tp and
mean direction broadcast (7 channels and a zero); targets the window means
of H, wave dissipation, roller dissipation, Fx, Fy.  Two families:

  ensemble  256 x 256 domains at dx = 4 m, bathymetries and sea states drawn
            as in the storm-scenario ensemble (the training configuration)
  site      the morphology site (1024 x 512 at dx = 3 m, barred rip-channel
            beach) under sea states drawn from the storm's range with seeds
            other than the storm run's, cut into 256 x 256 crops

In the site family the targets are made dimensionless with the offshore
wave height before the normalisation: H / Hm0 and the energy-like products
(Dw, DR, Fx, Fy) over Hm0^2 (``TARGET_POW``), so the network does not have
to learn the scale of the sea state; inference multiplies back
(``phys_norm``).  The ensemble family keeps per-channel-normalised
physical targets.  In the site
family the zs, uu, vv inputs of a sample are the means of an earlier window
(``ref_every``), matching the morphology run.  Samples stay in HBM as BF16
(NHWC), normalised per channel with the mean and standard deviation of the
first batch.  The site family is kept at full size
(1024 x 512: one sample per domain and window, the size the morphology run
infers on); the ensemble family is 256 x 256.

Training: Adam (lr 1e-3, batch 64) on the hand-written BF16 MFMA forward and
backward kernels (xbg_sur_*), 80/20 split, held-out loss reported.  torch
provides device buffers and the batch gathers only; every product of the
network and every update runs in the library's kernels.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

IN_CH = 8   # zb, zs, uu, vv, Hm0, Tp, dir, 0
OUT_CH = 5  # H, Dw, Dr, Fx, Fy
RAW_CH = 12
TARGET_POW = (1.0, 2.0, 2.0, 2.0, 2.0)  # targets divided by Hm0 ** pow before normalising


def phys_norm(norm10, target_pow, Hm0: float) -> np.ndarray:
    """Means and standard deviations (10 values) that turn the network's
    normalised outputs into physical H, Dw, DR, Fx, Fy for offshore Hm0."""
    n = np.asarray(norm10, np.float64)
    sc = np.float64(Hm0) ** np.asarray(target_pow, np.float64)
    return np.concatenate([n[:5] * sc, n[5:] * sc]).astype(np.float32)


def site_specs(n: int, seed: int = 99):
    """The morphology site under n sea states from the storm's range."""
    from .workloads import DomainSpec

    rng = np.random.default_rng(seed)
    out = []
    for d in range(n):
        Hm0 = float(rng.uniform(1.0, 4.0))
        out.append(DomainSpec(1024, 512, 3.0, 1 / 40.0, 1.0, Hm0, 8.0 + 6.0 * (Hm0 - 1.0) / 3.0,
                              float(rng.uniform(-20, 20)), 10.0, 24, 250e-6, 10_000 + seed * 100 + d))
    return out


def run_batch(specs, nwin: int, win_steps: int, spin_steps: int, on_window, dt_max: float = 0.16,
              ref_every: int = 0):
    """One ensemble of the given domains (all the same size): spin-up, then
    nwin windows, each window's samples (raw FP32 [ndom, ny, nx, 12] on the
    GPU, reused between windows) handed to on_window; returns (GPU seconds,
    host set-up seconds, cell-steps).  With ref_every = L the windows form
    segments of L: the first window of a segment is the reference, and the
    other L - 1 windows are samples whose zs, uu, vv inputs are the
    reference window's means (zb, the sea state and the targets their own),
    as the surrogate-assisted run feeds the network the last full-solver
    window's hydrodynamics."""
    import torch

    from . import Domain, workloads
    from .ensemble import Ensemble, RunSetup

    t0 = 50.0
    doms, runs, meta = [], [], []
    h0 = time.perf_counter()
    rtl = float(int(30 + (spin_steps + nwin * win_steps) * dt_max))
    for sp in specs:
        p, st, frc = workloads.build(sp, t0=t0, rtlength=rtl, zb_accum=0)
        d = Domain(p)
        for k, v in st.items():
            d.set(k, v)
        d.init_state(frc.Trep)
        d.set_bnd_series(frc.S, frc.Q)
        doms.append(d)
        runs.append(RunSetup(t0 + 1.0, 0.05, 1e9, t0, frc.dtbc, frc.Trep))
        meta.append((sp.Hm0, sp.Tp, math.radians(sp.dir_deg)))
    host_s = time.perf_counter() - h0
    ens = Ensemble(doms, runs)
    nfl = ens.window_enable()
    raw = torch.empty(nfl, dtype=torch.float32, device="cuda")
    ny, nx = specs[0].ny, specs[0].nx
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ens.step(spin_steps)
    ref = None
    for w in range(nwin):
        for _ in range(win_steps):
            ens.step(1)
            ens.window_add()
        ens.window_sample(win_steps, np.array(meta), raw.data_ptr())
        r = raw.view(len(specs), ny, nx, RAW_CH)
        if ref_every > 1:
            if w % ref_every == 0:
                ref = r[..., 1:4].clone()
                continue
            r = r.clone()
            r[..., 1:4] = ref
        on_window(r)
        # torch's reads of raw finish before the next window overwrites it
        torch.cuda.synchronize()
    ens.sync()
    gpu_s = time.perf_counter() - w0
    t, dt, steps, status = ens.status()
    if not (status == 0).all():
        raise RuntimeError(f"a training domain stopped: {status}")
    cell_steps = float(nx * ny * steps.sum())
    ens.close()
    for d in doms:
        d.close()
    return gpu_s, host_s, cell_steps


def crops(raw, size: int = 256):
    """Non-overlapping size x size crops of [S, ny, nx, C] samples."""
    S, ny, nx, C = raw.shape
    if ny == size and nx == size:
        return raw
    r = raw[:, : ny // size * size, : nx // size * size]
    r = r.reshape(S, ny // size, size, nx // size, size, C).permute(0, 1, 3, 2, 4, 5)
    return r.reshape(-1, size, size, C)


class Dataset:
    """Normalised BF16 samples in HBM; channel statistics from the first
    batch added.  scale: targets divided by Hm0 ** TARGET_POW first."""

    def __init__(self, scale: bool = False):
        self.X, self.T = [], []
        self.mean = self.std = None
        self.target_pow = np.asarray(TARGET_POW if scale else (0.0,) * OUT_CH, np.float32)

    def add(self, raw):
        import torch

        pw = torch.tensor(self.target_pow, dtype=raw.dtype, device=raw.device)
        tgt = raw[..., 7:] / raw[..., 4:5].clamp_min(0.1) ** pw
        if self.mean is None:
            fi = raw[..., :7].reshape(-1, 7)
            ft = tgt.reshape(-1, OUT_CH)
            self.mean = torch.cat([fi.mean(0), ft.mean(0)])
            self.std = torch.cat([fi.std(0), ft.std(0)]).clamp_min(1e-6)
        m, s = self.mean, self.std
        X = torch.zeros(raw.shape[:3] + (IN_CH,), dtype=torch.bfloat16, device=raw.device)
        X[..., :7] = ((raw[..., :7] - m[:7]) / s[:7]).to(torch.bfloat16)
        T = ((tgt - m[7:]) / s[7:]).to(torch.bfloat16).contiguous()
        self.X.append(X)
        self.T.append(T)

    def tensors(self):
        import torch

        X = torch.cat(self.X) if len(self.X) > 1 else self.X[0]
        T = torch.cat(self.T) if len(self.T) > 1 else self.T[0]
        self.X, self.T = [X], [T]
        return X, T

    def norm14(self):
        m, s = self.mean.cpu().numpy(), self.std.cpu().numpy()
        return np.concatenate([m[:7], s[:7]]).astype(np.float32)

    def norm10(self):
        m, s = self.mean.cpu().numpy(), self.std.cpu().numpy()
        return np.concatenate([m[7:], s[7:]]).astype(np.float32)


def train(X, T, epochs: int, batch: int = 64, lr: float = 1e-3, seed: int = 0, log=print, lr_final=None):
    """Adam on 80 % of the samples, at learning rate lr (with lr_final set,
    cosine decay from lr to lr_final over the run); returns (surrogate,
    per-epoch train loss, train seconds, held-out loss, steps)."""
    import torch

    from .surrogate import Surrogate

    S, H, W = X.shape[:3]
    g = torch.Generator().manual_seed(seed)
    perm = torch.randperm(S, generator=g)
    ntr = int(0.8 * S)
    tr, te = perm[:ntr].cuda(), perm[ntr:].cuda()
    sur = Surrogate(batch, H, W, seed=seed)
    hist = []
    w0 = time.perf_counter()
    nstep = 0
    total = epochs * ((ntr - batch) // batch + 1)
    lr_end = lr if lr_final is None else lr_final
    for ep in range(epochs):
        order = tr[torch.randperm(ntr, generator=g).cuda()]
        tot = 0.0
        nb = 0
        for k in range(0, ntr - batch + 1, batch):
            idx = order[k:k + batch]
            xb = X.index_select(0, idx).contiguous()
            tb = T.index_select(0, idx).contiguous()
            lr_k = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * nstep / max(total - 1, 1)))
            tot += sur.train_step(xb.data_ptr(), tb.data_ptr(), batch, lr_k)
            nb += 1
            nstep += 1
        sur.sync()
        hist.append(tot / max(nb, 1))
        log(f"epoch {ep}: train loss {hist[-1]:.5f}  ({time.perf_counter() - w0:.1f} s)")
    train_s = time.perf_counter() - w0
    held = evaluate(sur, X, T, te, batch)
    return sur, hist, train_s, held, nstep


def evaluate(sur, X, T, idx, batch):
    tot, n = 0.0, 0
    for k in range(0, len(idx), batch):
        ii = idx[k:k + batch]
        xb = X.index_select(0, ii).contiguous()
        tb = T.index_select(0, ii).contiguous()
        tot += sur.loss(xb.data_ptr(), tb.data_ptr(), len(ii)) * len(ii)
        n += len(ii)
    return tot / max(n, 1)


def main():
    import torch

    from . import workloads

    # torch's HIP runtime (the sample buffers) initialises before the
    # library opens the device
    torch.cuda.init()
    ap = argparse.ArgumentParser()
    ap.add_argument("--family", default="ensemble", choices=["ensemble", "site"])
    ap.add_argument("--batches", type=int, default=2)
    ap.add_argument("--domains", type=int, default=256)
    ap.add_argument("--windows", type=int, default=40)
    ap.add_argument("--win-steps", type=int, default=80)
    ap.add_argument("--spin-steps", type=int, default=600)
    ap.add_argument("--site-domains", type=int, default=16)
    ap.add_argument("--site-batches", type=int, default=4)
    ap.add_argument("--site-windows", type=int, default=40)
    ap.add_argument("--site-win-steps", type=int, default=130)
    ap.add_argument("--site-spin-steps", type=int, default=8000)
    ap.add_argument("--ref-every", type=int, default=10,
                    help="site family: zs, uu, vv inputs held from the first window of each segment of this many")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=None, help="default 64 (ensemble), 8 (site, full size)")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=None, help="cosine decay to this learning rate")
    ap.add_argument("--save", default=None, help="write parameters and normalisation (.npz)")
    a = ap.parse_args()
    ds = Dataset(scale=a.family == "site")
    info = {"gen_gpu_s": 0.0, "gen_host_s": 0.0, "gen_cell_steps": 0.0}

    def take(g, h, c):
        info["gen_gpu_s"] += g
        info["gen_host_s"] += h
        info["gen_cell_steps"] += c

    if a.family == "ensemble":
        batch = a.batch or 64
        for b in range(a.batches):
            specs = workloads.ensemble_specs(a.domains, seed=7000 + b, fixed_size=(256, 256, 4.0))
            take(*run_batch(specs, a.windows, a.win_steps, a.spin_steps, lambda r: ds.add(crops(r))))
            print(f"ensemble batch {b} done", flush=True)
    else:
        batch = a.batch or 8
        for b in range(a.site_batches):
            take(*run_batch(site_specs(a.site_domains, seed=99 + b), a.site_windows, a.site_win_steps,
                            a.site_spin_steps, ds.add, dt_max=0.1, ref_every=a.ref_every))
            print(f"site batch {b} done", flush=True)
    X, T = ds.tensors()
    print(f"{X.shape[0]} samples of {X.shape[1]} x {X.shape[2]}", flush=True)
    sur, hist, train_s, held, nstep = train(X, T, a.epochs, batch, a.lr, lr_final=a.lr_final)
    # forward 18.9 GFLOP per 256 x 256 image (the 17 convolutions); training
    # runs forward, data and weight gradients: about three times that
    flops = 3 * 18.9e9 * X.shape[1] * X.shape[2] / 65536.0 * batch * nstep
    out = {"samples": int(X.shape[0]), "family": a.family, "gen_gpu_s": info["gen_gpu_s"],
           "gen_host_s": info["gen_host_s"], "gen_cell_steps_per_s": info["gen_cell_steps"] / info["gen_gpu_s"],
           "train_s": train_s, "train_steps": nstep, "train_tflops_est": flops / train_s / 1e12,
           "train_loss": hist, "heldout_loss": held, "norm14": ds.norm14().tolist(), "norm10": ds.norm10().tolist()}
    if a.save:
        np.savez(a.save, params=sur.get_params(), norm14=ds.norm14(), norm10=ds.norm10(), heldout=held,
                 target_pow=ds.target_pow)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
