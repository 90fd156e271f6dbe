"""rabbitvar_amd: RabbitVar's per-region calling path on AMD MI355X (gfx950).

The compute path is the native library rabbitvar_amd/_build/librabbitvar_amd.so (HIP kernels + C++
host engine + C ABI, include/rabbitvar_amd.h).  This package holds the Python host mirror of that
interface (engine.Engine), the input readers (io), and the build script (build).
"""
from .io import BamFile, batch_from_records, read_fasta, read_rvb, write_rvb  # noqa: F401
