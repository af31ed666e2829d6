"""MI355X-native G3M^3 per-timestep multifluid hot path.

* ``include/g3m3.h``        -- the C-ABI (one entry point per reference
                               subroutine, Fortran by-reference convention)
* ``g3m3_amd/csrc``         -- the HIP kernels (gfx950) behind it
* ``g3m3_amd._lib``         -- ctypes binding of libg3m3.so (no fallback)
* ``g3m3_amd.driver.Run``   -- the reference's box-step / full-step
                               sequence over the C-ABI, data resident in HBM
* ``g3m3_amd.workload``     -- synthetic nested-box workloads
"""
__all__ = ["workload", "driver", "signatures"]
