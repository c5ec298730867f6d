"""Ring-polymer helpers (host side) for the PIMD integrator.

normal_mode_transform restates deepmd_jax/pimd.py:22-44: eigenvectors of the ring
Laplacian (2I - shift(+1) - shift(-1)), ascending eigenvalues, centroid eigenvalue
clamped to 0, frequencies P kT / hbar * sqrt(eigenvalue).  The device integrator
(md.hip: exact free-ring step and PILE-L O step) consumes C, the frequencies and the
per-mode Langevin factors.
"""
import numpy as np

from .constants import HBAR


def normal_mode_transform(n_bead, kT):
    ring_poly_freq = n_bead * kT / HBAR
    A = 2 * np.eye(n_bead) - np.roll(np.eye(n_bead), 1, axis=0) - np.roll(np.eye(n_bead), -1, axis=0)
    eva, eve = np.linalg.eigh(A)
    order = np.argsort(eva)
    eva, eve = eva[order], eve[:, order]
    if n_bead >= 2:
        eva[0] = 0.0
    return ring_poly_freq * np.sqrt(np.maximum(eva, 0.0)), eve


def pile_l_factors(freqs, dt, kT, n_bead, tau_t, pimd_lambda=1.0):
    """Per-mode O-step factors of PILE-L (deepmd_jax/pimd.py:92-111, md.py:341-351)."""
    gamma = 2.0 * pimd_lambda * np.asarray(freqs, np.float64)
    gamma[0] = 1.0 / tau_t
    c1 = np.exp(-gamma * dt)
    c2 = np.sqrt(np.maximum(kT * n_bead * (1.0 - c1 ** 2), 0.0))
    return gamma, c1, c2


def unwrap_across_beads(coord, box):
    """deepmd_jax/pimd.py:210-227: beads onto bead 0's periodic image, per axis for (3,)
    box lengths, through the fractional minimum image for a (3,3) cell."""
    coord = np.asarray(coord)
    ref = coord[0:1]
    box = np.asarray(box, np.float64)
    delta = coord - ref
    if box.size == 3:
        box = box.reshape(3)
        delta = delta - box * np.round(delta / box)
    else:
        cell = box.reshape(3, 3)
        frac = delta @ np.linalg.inv(cell)
        delta = (frac - np.round(frac)) @ cell
    return ref + delta


def centroid(coord, box):
    return unwrap_across_beads(coord, box).mean(axis=0)


def spring_energy(coord, mass, box, omega_p):
    """Ring-polymer spring energy 0.5 m omega_P^2 sum_b (x_b - x_{b-1})^2 over the beads
    unwrapped onto bead 0's image (deepmd_jax/pimd.py:82-89)."""
    xu = unwrap_across_beads(coord, box)
    dx = xu - np.roll(xu, 1, axis=0)
    return 0.5 * omega_p ** 2 * float((np.asarray(mass)[None, :, None] * dx ** 2).sum())
