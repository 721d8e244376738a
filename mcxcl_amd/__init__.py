"""mcxcl_amd - MI355X-native voxel Monte Carlo photon transport (MCX-CL compatible).

``run(cfg)`` / ``gpuinfo()`` / ``version()`` mirror the reference's ``pmcxcl``
package; the simulation itself runs in ``lib/libmcxhip.so`` (C ABI over HIP,
hand-written gfx950 kernels).  Importing raises ImportError if the native
library has not been built (``make -C mcxcl_amd``).
"""

from .pmcxcl import run, gpuinfo, version  # noqa: F401
from . import workloads, anchors  # noqa: F401

__all__ = ["run", "gpuinfo", "version", "workloads", "anchors"]

# mcxcl_amd.autograd (torch.autograd over the replay adjoint) is imported on demand:
# it needs torch, the rest of the package does not.
