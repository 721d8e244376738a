"""Differentiable detector signals: torch.autograd over the replay adjoint.

``detector_signals(mua, mus, cfg)`` runs the forward simulation on per-voxel
absorption ``mua`` and scattering ``mus`` fields (float32 [Nx, Ny, Nz], 1/mm;
voxels of label 0 in ``cfg['vol']`` stay background) and returns
S_d = sum of the exit weights of the photons detected by detector d, or, when
``cfg`` defines several sources (``srcpos`` with one row per source),
S_sd = the same sum over the photons launched by source s, shape [nsrc, ndet].
All sources run in one launch: each photon draws its source, and the kernel
records the source index with the detector id (src/mcx_core.cl:887-888).

The backward pass replays the detected photons from their saved RNG states, so
no trajectory is stored: each replayed photon retraces its path and deposits

  J(v)  = sum_i w_i L_iv        (outputtype 'jacobian')
  WP(v) = sum_i w_i n_iv        (outputtype 'wp': one deposit per scattering)

with L_iv the photon's path length in voxel v and n_iv its scattering count
there.  The upstream gradient g is folded into the replay weights (w_i g of
photon i's source and detector), so one replay of all detected photons gives
sum g_sd J_sd at once, and one more gives the WP term.

Absorption is continuous weight, w_i = exp(-sum_v mua_v L_iv), so
dS/dmua_v = -J(v).  The scattering derivative is the score of the path density
(free flights exp(-mus L), one factor mus per scattering event):
dS/dmus_v = WP(v)/mus_v - J(v).  The reference computes the same replay
deposits (src/mcx_core.cl:2568-2623, :2845-2858); the torch wrapper is new.
The forward run skips the fluence grid (issave2pt 0): detection does not
depend on it.
"""

import numpy as np
import torch

from . import pmcxcl

__all__ = ["DetectorSignals", "detector_signals", "signal_jacobians"]

# the simulation entry point (the HIP library through the pmcxcl binding)
_run = pmcxcl.run


def _as_numpy(t):
    return t.detach().to("cpu", torch.float32).contiguous().numpy()


def _signal_index(detword, ndet, nsrc):
    """flat (source, detector) index of each detected photon from its detid word"""
    word = detword.astype(np.int64)
    det = (word & 0xFFFF) - 1
    if nsrc == 1:
        return det
    return ((word >> 16) - 1) * ndet + det


def _nsrc(cfg):
    pos = np.atleast_2d(np.asarray(cfg["srcpos"], dtype=np.float64))
    return pos.shape[0]


class DetectorSignals(torch.autograd.Function):
    """S = DetectorSignals.apply(mua, mus, cfg) with cfg a pmcxcl dict holding a
    label ``vol``, ``prop`` (g and n per label), the source, ``detpos`` and the
    time window.  Runs on the device that the C library drives; the tensors may
    live on any torch device."""

    @staticmethod
    def forward(ctx, mua, mus, cfg):
        mm = np.stack([_as_numpy(mua), _as_numpy(mus)])
        nsrc = _nsrc(cfg)
        run_cfg = dict(cfg, muamus=mm, issavedet=1, issaveseed=1, savedetweight=1, issave2pt=0,
                       savedetflag=cfg.get("savedetflag", "DP"))
        if nsrc > 1:
            run_cfg["srcid"] = 0           # each photon draws its source; detid keeps it
        if "D" not in run_cfg["savedetflag"].upper() or "P" not in run_cfg["savedetflag"].upper():
            raise ValueError("detector_signals needs savedetflag with D (detector id) and P (path lengths)")
        out = _run(run_cfg)
        ndet = np.atleast_2d(np.asarray(cfg["detpos"])).shape[0]
        sums = np.zeros(nsrc * ndet, dtype=np.float64)
        if "detp" not in out:
            ctx.replay = None
        else:
            idx = _signal_index(out["detp"][0], ndet, nsrc)
            np.add.at(sums, idx, out["detw"].astype(np.float64))
            ctx.replay = (run_cfg, out["seeds"], out["detp"], out["detw"], idx)
        ctx.mus = mm[1]
        shape = (nsrc, ndet) if nsrc > 1 else (ndet,)
        return torch.as_tensor(sums.reshape(shape), dtype=torch.float64, device=mua.device)

    @staticmethod
    def backward(ctx, grad_s):
        shape = ctx.mus.shape
        if ctx.replay is None:
            z = torch.zeros(shape, dtype=torch.float32, device=grad_s.device)
            return z, z.clone(), None
        run_cfg, seeds, detp, detw, idx = ctx.replay
        g = grad_s.detach().to("cpu", torch.float64).numpy().ravel()
        # fold the upstream gradient into the replay weights: one replay of all
        # detected photons then deposits sum g_sd J_sd directly (replaydet 0)
        wg = (detw.astype(np.float64) * g[idx]).astype(np.float32)
        rep = dict(run_cfg, seed=seeds, detphotons=detp, replayweight=wg, replaydet=0, rawfield=1,
                   issaveseed=0, savedetweight=0, issave2pt=1)
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
    """Per-detector signals S_d(mua, mus), or S_sd with several sources, with
    gradients by replay (see module doc)."""
    return DetectorSignals.apply(mua, mus, cfg)


def signal_jacobians(mua, mus, cfg):
    """All source-detector signals and their full gradients at once.

    Returns (S, dS_dmua, dS_dmus) as numpy arrays: S [nsrc, ndet] (nsrc = 1 for
    one source) and the gradients [nsrc, ndet, Nx, Ny, Nz] float32.  One forward
    run saves the detected photons; two replays (outputtype 'jacobian' and
    'wp') with one output slice per detector (replaydet -1) and per source
    (srcid -1) give every pair's J_sd and WP_sd, the reference's per-detector
    replay Jacobian (src/mcx_utils.c:1355-1421, src/mcx_host.cpp:1397-1433)
    extended over sources.  The grids hold nsrc x ndet x Nvox float64 values
    on the device; each is converted to float32 there and read back once.  The
    gradients are views in the grids' memory order (x fastest), so they are not
    C-contiguous."""
    mua = np.ascontiguousarray(mua, dtype=np.float32)
    mus = np.ascontiguousarray(mus, dtype=np.float32)
    mm = np.stack([mua, mus])
    nsrc = _nsrc(cfg)
    ndet = np.atleast_2d(np.asarray(cfg["detpos"])).shape[0]
    run_cfg = dict(cfg, muamus=mm, issavedet=1, issaveseed=1, savedetweight=1, issave2pt=0,
                   savedetflag=cfg.get("savedetflag", "DP"))
    if nsrc > 1:
        run_cfg["srcid"] = 0
    out = _run(run_cfg)
    shape = mua.shape
    sums = np.zeros(nsrc * ndet, dtype=np.float64)
    if "detp" not in out:
        z = np.zeros((nsrc, ndet) + shape, np.float32)
        return sums.reshape(nsrc, ndet), z, z.copy()
    idx = _signal_index(out["detp"][0], ndet, nsrc)
    np.add.at(sums, idx, out["detw"].astype(np.float64))
    rep = dict(run_cfg, seed=out["seeds"], detphotons=out["detp"], replayweight=out["detw"], replaydet=-1,
               rawfield=1, issaveseed=0, savedetweight=0, issave2pt=1)
    if nsrc > 1:
        rep["srcid"] = -1
    if ndet == 1:
        rep["replaydet"] = 0
    unit = float(run_cfg.get("unitinmm", 1.0))

    rev = shape[::-1]

    def slices(outputtype, scale):
        """[nsrc, ndet, Nz, Ny, Nx] in memory order, times scale"""
        f = _run(dict(rep, outputtype=outputtype, rawfield=0, rawfield32=1, rawonly=1))["flux32"]
        m = f.reshape(-1, order="F").reshape((nsrc, ndet, -1) + rev)   # memory order: src, det, gate, z, y, x
        m = m[:, :, 0] if m.shape[2] == 1 else m.sum(2)
        m *= np.float32(scale)
        return m

    dmua = slices("jacobian", -unit)                     # -J
    dmus = slices("wp", 1.0)
    with np.errstate(divide="ignore"):
        inv = np.where(mus > 0, 1.0 / mus, 0.0).astype(np.float32)
    dmus *= np.ascontiguousarray(inv.transpose(2, 1, 0))  # WP / mus
    dmus += dmua                                         # - J
    view = (0, 1, 4, 3, 2)                               # -> [nsrc, ndet, Nx, Ny, Nz]
    return sums.reshape(nsrc, ndet), dmua.transpose(view), dmus.transpose(view)
