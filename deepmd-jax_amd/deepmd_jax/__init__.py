"""deepmd-jax's Deep Potential se_a energy/force path and MD step on MI355X (gfx950).

Public surface mirroring the reference package:
  deepmd_jax.md.Simulation, deepmd_jax.train.evaluate,
  deepmd_jax.utils.load_model / save_model.
All compute runs in libdpmd_hip.so (hand-written HIP kernels) through deepmd_jax._lib.
"""
__version__ = '0.1.0'
