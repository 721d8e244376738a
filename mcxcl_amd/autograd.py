"""Differentiable detector signals: torch.autograd over the replay adjoint.

``detector_signals(mua, mus, cfg)`` runs the forward simulation on per-voxel
absorption ``mua`` and scattering ``mus`` fields (float32 [Nx, Ny, Nz], 1/mm;
voxels of label 0 in ``cfg['vol']`` stay background) and returns
S_d = sum of the exit weights of the photons detected by detector d.

The backward pass replays the detected photons from their saved RNG states, so
no trajectory is stored: each replayed photon retraces its path and deposits

  J_d(v)  = sum_{i in d} w_i L_iv        (outputtype 'jacobian')
  WP_d(v) = sum_{i in d} w_i n_iv        (outputtype 'wp': one deposit per scattering)

The upstream gradient g_d is folded into the replay weights (w_i g_d(i)), so one
replay of all detected photons deposits sum_d g_d J_d at once.

with L_iv the photon's path length in voxel v and n_iv its scattering count
there.  Absorption is continuous weight, w_i = exp(-sum_v mua_v L_iv), so
dS_d/dmua_v = -J_d(v).  The scattering derivative is the score of the path
density (free flights exp(-mus L), one factor mus per scattering event):
dS_d/dmus_v = WP_d(v)/mus_v - J_d(v).  The reference computes the same replay
deposits (src/mcx_core.cl:2568-2623, :2845-2858); the torch wrapper is new.
"""

import numpy as np
import torch

from . import pmcxcl

__all__ = ["DetectorSignals", "detector_signals"]

# the simulation entry point (the HIP library through the pmcxcl binding)
_run = pmcxcl.run


def _as_numpy(t):
    return t.detach().to("cpu", torch.float32).contiguous().numpy()


def _detector_sums(detid, detw, ndet):
    s = np.zeros(ndet, dtype=np.float64)
    np.add.at(s, detid - 1, detw.astype(np.float64))
    return s


class DetectorSignals(torch.autograd.Function):
    """S = DetectorSignals.apply(mua, mus, cfg) with cfg a pmcxcl dict holding a
    label ``vol``, ``prop`` (g and n per label), the source, ``detpos`` and the
    time window.  Runs on the device that the C library drives; the tensors may
    live on any torch device."""

    @staticmethod
    def forward(ctx, mua, mus, cfg):
        mm = np.stack([_as_numpy(mua), _as_numpy(mus)])
        run_cfg = dict(cfg, muamus=mm, issavedet=1, issaveseed=1, savedetweight=1,
                       savedetflag=cfg.get("savedetflag", "DP"))
        if "D" not in run_cfg["savedetflag"].upper() or "P" not in run_cfg["savedetflag"].upper():
            raise ValueError("detector_signals needs savedetflag with D (detector id) and P (path lengths)")
        out = _run(run_cfg)
        ndet = np.atleast_2d(np.asarray(cfg["detpos"])).shape[0]
        if "detp" not in out:
            sums = np.zeros(ndet)
            ctx.replay = None
        else:
            detid = out["detp"][0].astype(np.int64) & 0xFFFF
            sums = _detector_sums(detid, out["detw"], ndet)
            ctx.replay = (run_cfg, out["seeds"], out["detp"], out["detw"])
        ctx.mus = mm[1]
        ctx.ndet = ndet
        return torch.as_tensor(sums, dtype=torch.float64, device=mua.device)

    @staticmethod
    def backward(ctx, grad_s):
        shape = ctx.mus.shape
        if ctx.replay is None:
            z = torch.zeros(shape, dtype=torch.float32, device=grad_s.device)
            return z, z.clone(), None
        run_cfg, seeds, detp, detw = ctx.replay
        g = grad_s.detach().to("cpu", torch.float64).numpy()
        # fold the upstream gradient into the replay weights: one replay of all
        # detectors then deposits sum_d g_d J_d directly (replaydet 0)
        detid = detp[0].astype(np.int64) & 0xFFFF
        wg = (detw.astype(np.float64) * g[detid - 1]).astype(np.float32)
        rep = dict(run_cfg, seed=seeds, detphotons=detp, replayweight=wg, replaydet=0, rawfield=1,
                   issaveseed=0, savedetweight=0)
        unit = float(run_cfg.get("unitinmm", 1.0))
        gJ = _run(dict(rep, outputtype="jacobian"))["flux64"].sum(-1).reshape(shape) * unit
        gW = _run(dict(rep, outputtype="wp"))["flux64"].sum(-1).reshape(shape)
        # J holds path lengths in grid units; per-mm derivatives carry unitinmm
        mus = ctx.mus.astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            gmus = np.where(mus > 0, gW / mus, 0.0) - gJ
        grad_mua = torch.as_tensor(-gJ, dtype=torch.float32, device=grad_s.device)
        grad_mus = torch.as_tensor(gmus, dtype=torch.float32, device=grad_s.device)
        return grad_mua, grad_mus, None


def detector_signals(mua, mus, cfg):
    """Per-detector signals S_d(mua, mus) with gradients by replay (see module doc)."""
    return DetectorSignals.apply(mua, mus, cfg)
