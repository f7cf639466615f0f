"""mrm_amd: MI355X (gfx950) HIP implementation of mrm's granular-media time-stepping path.

The compute path is the C-ABI library libmrmg.so built from mrm_amd/csrc (declared in
include/mrmg.h).  This package holds its Python host side: the ctypes binding (abi), the
mirror of the reference's llg_m driver interface (llg), and synthetic media (media).
"""
__all__ = ["abi", "llg", "media"]
