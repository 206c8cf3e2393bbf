"""Surrogate-assisted long-horizon morphology and the plain GPU solver on the
same forcing.

Site: 1024 x 512 cells at dx = 3 m, the barred rip-channel beach of the
large configuration, 24 direction bins, morfac 10, FP64 bed accumulation.
Forcing: a 24 h storm in morphological time (Hm0 1 -> 4 -> 1 m as a Gaussian
of sigma 6 h centred at 12 h; Tp = 8 + 6 (Hm0 - 1)/3 s; mean direction from
-20 to +20 deg), i.e. 24 h / morfac of hydrodynamic time, as seeded JONSWAP
wave-group blocks (wavebndtype 4 with one record per block, regenerated per
block as the reference regenerates per record, Wavestep.cu:512-753).

Surrogate-assisted run: windows of one peak period; after a full-solver
spin-up, each window's wave products come from the surrogate applied to the
previous window's means (zb, zs, uu, vv) and the offshore sea state, and the
flow, sediment and bed steps run without the wave action solve
(xbg_step_nowave).  Every ``ckpt_every`` windows the full solver runs a
re-spin window and a checkpoint window whose mean H judges the surrogate's
prediction for it; an error above ``tol`` (relative L2 over wet cells) puts
the run on the full solver, window by window, until the surrogate's
prediction is within tol again.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import time

import numpy as np

MORFAC = 10.0
STORM_S = 24 * 3600.0


def storm_state(t_morph: float):
    """Hm0 (m), Tp (s), mean direction (deg) at morphological time t."""
    Hm0 = 1.0 + 3.0 * math.exp(-0.5 * ((t_morph - 12 * 3600.0) / (6 * 3600.0)) ** 2)
    Tp = 8.0 + 6.0 * (Hm0 - 1.0) / 3.0
    return Hm0, Tp, -20.0 + 40.0 * t_morph / STORM_S


def site_params(hydro_s: float):
    from . import setups

    return setups.large_params(nx=1024, ny=512, dx=3.0, morfac=MORFAC, endtime=hydro_s, zb_accum=2)


class StormBlocks:
    """JONSWAP wave-group blocks of block_s hydrodynamic seconds, each with
    the storm's sea state at the block's middle (seed = block index + 1000
    seed)."""

    def __init__(self, p, hydro_s: float, block_s: float = 600.0, offdepth: float = 20.0, seed: int = 0):
        from . import setups

        self.block_s = block_s
        self.blocks = []
        self.state = []
        nb = int(math.ceil(hydro_s / block_s))
        for b in range(nb):
            Hm0, Tp, d = storm_state((b + 0.5) * block_s * MORFAC)
            self.state.append((Hm0, Tp, d))
            self.blocks.append(setups.JonswapForcing(p, Hm0, Tp, dir_deg=d, s=10.0, seed=1000 * seed + b,
                                                     offdepth=offdepth, rtlength=block_s + 10.0, t0=b * block_s))

    def index(self, t: float) -> int:
        return min(int(t // self.block_s), len(self.blocks) - 1)


def new_domain(p, blocks):
    from . import Domain, setups

    elev = setups.barred_bathy(p.nx, p.ny, p.dx)
    st = setups.initial_state(p, elev)
    d = Domain(p)
    for k, v in st.items():
        d.set(k, v)
    d.init_state(blocks.blocks[0].Trep)
    return d


def run_plain(p, blocks, hydro_s: float):
    d = new_domain(p, blocks)
    zb0 = d.get("zb").astype(np.float64)
    t, dt, steps = 0.0, 0.05, 0
    w0 = time.perf_counter()
    for b, f in enumerate(blocks.blocks):
        if t >= hydro_s:
            break
        d.set_bnd_series(f.S, f.Q)
        t, dt, n = d.advance_window(1, f.t0, f.dtbc, f.Trep, min(f.t0 + blocks.block_s, hydro_s), hydro_s, False,
                                    t, dt)
        steps += n
    d.sync()
    wall = time.perf_counter() - w0
    zb = d.get("zb").astype(np.float64)
    d.close()
    return {"wall_s": wall, "steps": steps, "t": t}, zb0, zb


def run_surrogate(p, blocks, hydro_s: float, ckpt: dict, spin_windows: int = 40, ckpt_every: int = 50,
                  tol: float = 0.05, log=print):
    import torch

    from .surrogate import Surrogate

    d = new_domain(p, blocks)
    zb0 = d.get("zb").astype(np.float64)
    ny, nx = p.ny, p.nx
    sur = Surrogate(1, ny, nx, seed=0)
    sur.set_params(ckpt["params"])
    norm14 = np.asarray(ckpt["norm14"], np.float32)
    norm10 = np.asarray(ckpt["norm10"], np.float32)
    x8 = torch.empty(ny * nx * 8, dtype=torch.bfloat16, device="cuda")
    y = torch.empty(ny * nx * 5, dtype=torch.float32, device="cuda")
    L = d.L
    ptr = {k: d.device_ptr(k) for k in ("zb", "zsmean", "uumean", "vvmean")}
    t, dt = 0.0, 0.05
    w, cur_block = 0, -1
    have_means, fallback = False, False
    stats = {"full_windows": 0, "sur_windows": 0, "steps_full": 0, "steps_sur": 0, "checkpoints": [],
             "fallback_windows": 0}
    w0 = time.perf_counter()

    def predict(Hm0, Tp, dgrid):
        rc = L.xbg_sur_pack_inputs(ptr["zb"], ptr["zsmean"], ptr["uumean"], ptr["vvmean"], nx * ny, Hm0, Tp, dgrid,
                                   norm14.ctypes.data_as(C.POINTER(C.c_float)), x8.data_ptr(), d.stream)
        if rc != 0:
            raise RuntimeError("xbg_sur_pack_inputs failed")
        sur.forward(x8.data_ptr(), 1, y.data_ptr())

    def window(mode, t_end):
        nonlocal t, dt
        d.reset_means()
        f = blocks.blocks[cur_block]
        t, dt, n = d.advance_window(mode, f.t0, f.dtbc, f.Trep, t_end, hydro_s, True, t, dt)
        d.divide_means(max(n, 1))
        key = "full" if mode else "sur"
        stats[key + "_windows"] += 1
        stats["steps_" + key] += n
        return n

    while t < hydro_s:
        b = blocks.index(t)
        if b != cur_block:
            cur_block = b
            d.set_bnd_series(blocks.blocks[b].S, blocks.blocks[b].Q)
        Hm0, Tp, ddeg = blocks.state[b]
        dgrid = math.radians(ddeg)
        t_end = min(t + Tp, (b + 1) * blocks.block_s, hydro_s)
        warm = have_means and w >= spin_windows
        checkpoint = warm and (w % ckpt_every == 0)
        if warm and not checkpoint and not fallback:
            predict(Hm0, Tp, dgrid)
            d.set_wave_products(y.data_ptr(), norm10, dgrid)
            window(0, t_end)
        elif warm:
            if checkpoint:
                window(1, t_end)   # the wave field catches up with the bed and currents
                if t >= hydro_s:
                    break
                t_end = min(t + Tp, (b + 1) * blocks.block_s, hydro_s)
            predict(Hm0, Tp, dgrid)
            window(1, t_end)
            Hp = y.view(ny, nx, 5)[..., 0].double() * float(norm10[5]) + float(norm10[0])
            Hp = Hp.clamp_min(0.0).cpu().numpy()
            Hf = d.get("Hmean").astype(np.float64)
            wet = d.get("hh") > p.eps
            err = float(np.linalg.norm((Hp - Hf)[wet]) / max(np.linalg.norm(Hf[wet]), 1e-30))
            stats["checkpoints"].append((w, round(t, 2), err))
            if fallback:
                stats["fallback_windows"] += 1
            fallback = err > tol
            log(f"window {w} t={t:.0f} s checkpoint H error {err:.3f} -> {'full solver' if fallback else 'surrogate'}")
        else:
            window(1, t_end)
        have_means = True
        w += 1
    d.sync()
    wall = time.perf_counter() - w0
    zb = d.get("zb").astype(np.float64)
    sur.close()
    d.close()
    stats.update({"wall_s": wall, "windows": w, "t": t})
    return stats, zb0, zb


def compare(p, zb0, zb_plain, zb_sur):
    area = p.dx * p.dx
    dz = zb_sur - zb_plain
    return {"zb_rmse_m": float(np.sqrt(np.mean(dz ** 2))), "zb_maxabs_m": float(np.abs(dz).max()),
            "volume_change_plain_m3": float((zb_plain - zb0).sum() * area),
            "volume_change_surrogate_m3": float((zb_sur - zb0).sum() * area),
            "bed_change_rms_plain_m": float(np.sqrt(np.mean((zb_plain - zb0) ** 2)))}


def main():
    import torch

    # torch's HIP runtime (device buffers of the surrogate's inputs and
    # outputs) initialises before the library opens the device
    torch.cuda.init()
    ap = argparse.ArgumentParser()
    ap.add_argument("--ckpt", required=True, help=".npz from sur_train --save")
    ap.add_argument("--hours", type=float, default=24.0, help="morphological hours of the storm to run")
    ap.add_argument("--ckpt-every", type=int, default=50)
    ap.add_argument("--tol", type=float, default=0.05)
    ap.add_argument("--save", default=None, help="write the initial and final bed levels (.npz)")
    a = ap.parse_args()
    hydro_s = a.hours * 3600.0 / MORFAC
    p = site_params(hydro_s)
    blocks = StormBlocks(p, hydro_s)
    ck = dict(np.load(a.ckpt))
    s, zb0, zs_ = run_surrogate(p, blocks, hydro_s, ck, ckpt_every=a.ckpt_every, tol=a.tol)
    pl, zb0b, zp = run_plain(p, blocks, hydro_s)
    out = {"hydro_s": hydro_s, "plain": pl, "surrogate": {k: v for k, v in s.items() if k != "checkpoints"},
           "checkpoints": s["checkpoints"], "speedup": pl["wall_s"] / s["wall_s"]}
    out.update(compare(p, zb0, zp, zs_))
    if a.save:
        np.savez_compressed(a.save, zb0=zb0.astype(np.float32), zb_plain=zp.astype(np.float32),
                            zb_surrogate=zs_.astype(np.float32))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
