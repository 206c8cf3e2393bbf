"""Surrogate-assisted long-horizon morphology and the plain GPU solver on the
same forcing.

Site: 1024 x 512 cells at dx = 3 m, the barred rip-channel beach of the
large configuration, 24 direction bins, morfac 10, FP64 bed accumulation.
Forcing: a 24 h storm in morphological time (Hm0 1 -> 4 -> 1 m as a Gaussian
of sigma 6 h centred at 12 h; Tp = 8 + 6 (Hm0 - 1)/3 s; mean direction from
-20 to +20 deg), i.e. 24 h / morfac of hydrodynamic time, as seeded JONSWAP
wave-group blocks (wavebndtype 4 with one record per block, regenerated per
block as the reference regenerates per record, Wavestep.cu:512-753).

Surrogate-assisted run: windows of one peak period.  The run starts on the
full solver (spin-up of ``spin_s`` hydrodynamic seconds, the time the wave
field needs to cross the domain), then judges the surrogate: for ``check``
full-solver windows the surrogate's prediction of each window is kept
beside the window's mean H.  The network's inputs are the current offshore
sea state and the zb, zs, uu, vv of the last full-solver window (the bed at
its end, the others its means): windows run on the surrogate's products
never feed back into its inputs.  They would close a loop in which the
surrogate's errors drive currents and bed changes that drive its next
prediction further out of its training range.  Wave groups make the mean H of single
windows differ by 15-25 % from their mean over many windows, so a single
window cannot judge a mean-field surrogate.  The checkpoint statistic is the
relative L2 error (wet cells) of the ``check``-window mean prediction P
against the ``check``-window mean H of the solver F with the wave-group
noise of F removed:

    e^2 = (|P - F|^2 - sum over cells of s^2 / K) / |F|^2

where s^2 is each cell's sample variance of the K window means (the
expected |F - <F>|^2 of a K-window mean is the sum of the per-cell
variances over K); the raw error is reported beside it.  Within ``tol``
the surrogate takes over: each window's
wave products come from it and the flow, sediment and bed steps run without
the wave action solve (xbg_step_nowave).  After ``ckpt_every`` surrogate
windows the full solver runs ``respin`` windows (its wave field catches up
with the bed, currents and sea state) and a checkpoint of ``check`` windows
judged as above; above ``tol`` the run stays on the full solver, checkpoint
after checkpoint, until the surrogate is within ``tol`` again.

Stability guard: when the surrogate takes over, the whole device state is
saved (xbg_state_save).  A surrogate window that crashes (non-finite time
step) or ends with a current faster than ``vfac`` times the fastest of the
checkpoint windows sends the run back to that state (xbg_state_restore) and
onto the full solver for ``ckpt_every`` windows, then to a new checkpoint.
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


def checkpoint_error(P, F, F2, K: int, wet):
    """(noise-corrected, raw) relative L2 error of the mean prediction P
    against the K-window mean F of the solver's H, F2 the mean of the
    squares; wet selects the cells."""
    import torch

    num = float(((P - F)[wet] ** 2).sum())
    den = max(float((F[wet] ** 2).sum()), 1e-30)
    var = (F2 - F * F).clamp_min(0.0) * (K / max(K - 1, 1))  # per-cell sample variance of the window means
    noise = float(var[wet].sum()) / K
    return math.sqrt(max(num - noise, 0.0) / den), math.sqrt(num / den)


def run_surrogate(p, blocks, hydro_s: float, ckpt: dict, spin_s: float = 600.0, ckpt_every: int = 50,
                  respin: int = 4, check: int = 12, tol: float = 0.05, vfac: float = 2.0, log=print):
    import torch

    from .domain import ModelCrashed
    from .sur_train import phys_norm
    from .surrogate import Surrogate

    d = new_domain(p, blocks)
    zb0 = d.get("zb").astype(np.float64)
    ny, nx = p.ny, p.nx
    sur = Surrogate(1, ny, nx, seed=0)
    sur.set_params(ckpt["params"])
    norm14 = np.asarray(ckpt["norm14"], np.float32)
    norm10 = np.asarray(ckpt["norm10"], np.float32)
    tpow = np.asarray(ckpt.get("target_pow", np.zeros(5)), np.float64)
    x8 = torch.empty(ny * nx * 8, dtype=torch.bfloat16, device="cuda")
    y = torch.empty(ny * nx * 5, dtype=torch.float32, device="cuda")
    L = d.L
    # zb and the zs, uu, vv window means of the last full-solver window
    ref = {k: torch.zeros(ny * nx, dtype=torch.float32, device="cuda") for k in ("zb", "zsmean", "uumean", "vvmean")}
    t, dt = 0.0, 0.05
    cur_block = -1
    stats = {"full_windows": 0, "sur_windows": 0, "steps_full": 0, "steps_sur": 0, "checkpoints": [],
             "fallback_checks": 0, "rollbacks": [], "sur_windows_kept": 0}
    w0 = time.perf_counter()

    def predict(Hm0, Tp, dgrid):
        """The surrogate's outputs for the coming window (in y) and the
        physical normalisation for this sea state."""
        rc = L.xbg_sur_pack_inputs(ref["zb"].data_ptr(), ref["zsmean"].data_ptr(), ref["uumean"].data_ptr(),
                                   ref["vvmean"].data_ptr(), nx * ny, Hm0, Tp, dgrid,
                                   norm14.ctypes.data_as(C.POINTER(C.c_float)), x8.data_ptr(), d.stream)
        if rc != 0:
            raise RuntimeError("xbg_sur_pack_inputs failed")
        sur.forward(x8.data_ptr(), 1, y.data_ptr())
        return phys_norm(norm10, tpow, Hm0)

    def window(mode, t_end):
        nonlocal t, dt
        d.reset_means()
        f = blocks.blocks[cur_block]
        t, dt, n = d.advance_window(mode, f.t0, f.dtbc, f.Trep, t_end, hydro_s, True, t, dt)
        d.divide_means(max(n, 1))
        if mode:
            for k, v in ref.items():
                v.copy_(torch.from_numpy(d.get(k).reshape(-1)))
        key = "full" if mode else "sur"
        stats[key + "_windows"] += 1
        stats["steps_" + key] += n
        return n

    def vmax():
        return float(max(np.abs(d.get("uu")).max(), np.abs(d.get("vv")).max()))

    phase, count = "spin", 0
    pred_sum = full_sum = full_sq = None
    saved = None      # (t, dt, block) of the state saved when the surrogate took over
    v_check = 0.0     # fastest current at the end of the checkpoint windows
    while t < hydro_s:
        b = blocks.index(t)
        if b != cur_block:
            cur_block = b
            d.set_bnd_series(blocks.blocks[b].S, blocks.blocks[b].Q)
        Hm0, Tp, ddeg = blocks.state[b]
        dgrid = math.radians(ddeg)
        t_end = min(t + Tp, (b + 1) * blocks.block_s, hydro_s)
        if phase == "spin":
            window(1, t_end)
            if t >= spin_s:
                phase, count = "check", 0
        elif phase == "sur":
            nm = predict(Hm0, Tp, dgrid)
            d.set_wave_products(y.data_ptr(), nm, dgrid)
            try:
                window(0, t_end)
                v = vmax()
                bad = not (v <= vfac * v_check)
            except ModelCrashed:
                v, bad = float("nan"), True
            count += 1
            if bad:
                d.state_restore()
                t, dt, cur_block = saved
                d.set_bnd_series(blocks.blocks[cur_block].S, blocks.blocks[cur_block].Q)
                stats["rollbacks"].append((round(t, 2), count, v))
                log(f"t={t:.0f} s: surrogate window {count} unstable (max current {v:.2f} m/s, checkpoint "
                    f"{v_check:.2f}) -> back to the checkpoint, full solver")
                phase, count = "fallback", 0
            elif count >= ckpt_every:
                stats["sur_windows_kept"] += count
                phase, count = "respin", 0
        elif phase in ("respin", "fallback"):
            window(1, t_end)
            count += 1
            if count >= (respin if phase == "respin" else ckpt_every):
                phase, count = "check", 0
        else:  # check: full windows with the surrogate's predictions beside them
            nm = predict(Hm0, Tp, dgrid)
            window(1, t_end)
            hh = torch.from_numpy(d.get("hh")).cuda()
            wet = hh > p.eps
            Hp = (y.view(ny, nx, 5)[..., 0].double() * float(nm[5]) + float(nm[0])).clamp_min(0.0)
            Hp = torch.minimum(Hp, p.gammax * hh.double()) * wet
            Hf = torch.from_numpy(d.get("Hmean")).cuda().double()
            pred_sum = Hp if count == 0 else pred_sum + Hp
            full_sum = Hf if count == 0 else full_sum + Hf
            full_sq = Hf * Hf if count == 0 else full_sq + Hf * Hf
            v_check = vmax() if count == 0 else max(v_check, vmax())
            count += 1
            if count >= check:
                err, raw = checkpoint_error(pred_sum / check, full_sum / check, full_sq / check, check, wet)
                ok = err <= tol
                stats["checkpoints"].append((round(t, 2), round(err, 4), round(raw, 4), bool(ok)))
                log(f"t={t:.0f} s ({t * MORFAC / 3600.0:.2f} h morph.) checkpoint H error {err:.3f} "
                    f"(raw {raw:.3f}) over {check} windows -> {'surrogate' if ok else 'full solver'}")
                if ok:
                    d.state_save()
                    saved = (t, dt, cur_block)
                    phase, count = "sur", 0
                else:
                    stats["fallback_checks"] += 1
                    count = 0
    if phase == "sur":
        stats["sur_windows_kept"] += count
    d.sync()
    wall = time.perf_counter() - w0
    zb = d.get("zb").astype(np.float64)
    sur.close()
    d.close()
    stats.update({"wall_s": wall, "t": t})
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
    ap.add_argument("--respin", type=int, default=4)
    ap.add_argument("--check", type=int, default=12)
    ap.add_argument("--spin", type=float, default=600.0, help="full-solver spin-up (hydrodynamic s)")
    ap.add_argument("--tol", type=float, default=0.05)
    ap.add_argument("--vfac", type=float, default=2.0, help="stability bound on the current, times the checkpoint's")
    ap.add_argument("--save", default=None, help="write the initial and final bed levels (.npz)")
    a = ap.parse_args()
    hydro_s = a.hours * 3600.0 / MORFAC
    p = site_params(hydro_s)
    blocks = StormBlocks(p, hydro_s)
    ck = dict(np.load(a.ckpt))
    s, zb0, zs_ = run_surrogate(p, blocks, hydro_s, ck, spin_s=a.spin, ckpt_every=a.ckpt_every, respin=a.respin,
                                check=a.check, tol=a.tol, vfac=a.vfac)
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
