"""MI355X-native SU(2) DMRG hot path (H_eff matvec, sector truncation,
environment update) behind the C ABI of ``include/su2dmrg.h``.

The compute lives in ``libsu2dmrg.so`` (HIP for gfx950, built from
``csrc/`` by ``su2dmrg_amd.build``); ``su2dmrg_amd.lib`` binds it.
"""
__all__ = ["build", "lib", "synthetic"]
