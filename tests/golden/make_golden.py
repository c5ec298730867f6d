"""Writes tests/golden/water64_seed0.npz: the 64-molecule seed-0 box, the seed-0
model's normalisation statistics and the oracle's FP64 energy, forces and virial.
These are regression pins of the oracle and the generators (the reference ships no
fixtures; its JAX code cannot run here)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, 'deepmd-jax_amd'))
sys.path.insert(0, ROOT)

from deepmd_jax import synthetic as syn  # noqa: E402
from oracle import dp_oracle as O  # noqa: E402

coord, box, types = syn.water_box(64, seed=0)
params, var = syn.init_model(0)
E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
np.savez(os.path.join(HERE, 'water64_seed0.npz'), coord=coord, box=box, types=types, energy=E.item(),
         force=F.numpy(), virial=W.numpy(), sr_mean=params['sr_mean'], sr_std=params['sr_std'],
         nnbrs=params['Nnbrs'])
print('E =', E.item())
