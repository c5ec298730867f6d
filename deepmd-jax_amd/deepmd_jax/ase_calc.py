"""ASE calculator on the GPU engine: DPJaxCalculator of deepmd_jax/ase_calc.py:18-144.

Like the reference module this one imports ASE at the top, so it needs ASE installed.
Energies, forces and the stress (Voigt, ASE sign convention, eV/A^3) come from
deepmd_jax.evaluator.energy_forces_stress; orthorhombic and general (triclinic) cells, and
cells narrower than 2 rcut are evaluated as periodic supercells (the reference's
lattice-image path).
"""
import numpy as np
from ase.calculators.calculator import Calculator, all_changes

from .evaluator import Evaluator, energy_forces_stress
from .utils import load_model


class DPJaxCalculator(Calculator):
    implemented_properties = ['energy', 'forces', 'stress']

    def __init__(self, model_path, type_idx=None, dtype=None, precision='fp32', device=0, **kwargs):
        Calculator.__init__(self, **kwargs)
        self._model, self._variables = load_model(model_path)
        type_idx_np = np.array(type_idx).astype(int)
        ct = self._model.params.get('chemical_types')
        if ct is not None:
            unknown = set(type_idx_np.tolist()) - set(ct)
            if unknown:
                raise ValueError('Atomic numbers %s in type_idx are not in model.params["chemical_types"]=%s'
                                 % (sorted(unknown), ct))
            z_to_idx = {z: i for i, z in enumerate(ct)}
            type_idx_np = np.array([z_to_idx[z] for z in type_idx_np], dtype=int)
        self._type_idx = type_idx_np
        self._precision = precision
        self._evaluator = Evaluator(self._model.params, self._variables, device)
        print('# Initializing the DPJaxCalculator')

    def calculate(self, atoms=None, properties=None, system_changes=all_changes):
        Calculator.calculate(self, atoms, properties, system_changes)
        cell = np.asarray(self.atoms.get_cell(complete=True))
        e, f, s = energy_forces_stress(self._evaluator, self.atoms.get_positions(), cell, self._type_idx,
                                       self._precision)
        self.results['energy'] = e
        self.results['forces'] = np.asarray(f)
        self.results['stress'] = s
