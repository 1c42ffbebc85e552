"""MI355X-native forced-DNS + droplet RK4 step (hot path of Rajarshi-prime/lucky-droplets).

Kernels live in libld.so (HIP, gfx950; C-ABI in include/ld.h).  Python keeps
the reference's function signatures:

  lucky_droplets_amd.dns.ForcedDNS          rfft_mpi, irfft_mpi, forcing, clip_zero,
                                            full_RHS, RK4, step   (forced-dns-sm.py)
  lucky_droplets_amd.particles.MPI_particles  pRHS, interp_cosine, exterp_cosine_scalar,
                                            send, update_intrinsic (particles.py)
"""
from .config import Config, preset  # noqa: F401

__all__ = ["Config", "preset"]
