"""ctypes glue of the C++/OpenMP CPU restatement (oracle/dp_cpu.cpp) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
build() compiles oracle/_build/libdp_cpu.so with g++ (-O3, OpenMP, x86-64-v3); energy_force()
evaluates one orthorhombic frame with a neighbour list from dp_oracle.typed_nlist, in FP64 (the
independent pin of the torch oracle) or FP32 (the CPU baseline).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'dp_cpu.cpp')
LIB = os.path.join(HERE, '_build', 'libdp_cpu.so')
_lib = None


class _Model(ctypes.Structure):
    _fields_ = [('ntypes', ctypes.c_int32), ('axis', ctypes.c_int32), ('n_embed', ctypes.c_int32),
                ('n_fit', ctypes.c_int32), ('embed_w', ctypes.c_int32 * 8), ('fit_w', ctypes.c_int32 * 8),
                ('rcut', ctypes.c_double), ('nnbrs', ctypes.c_double), ('out_norm', ctypes.c_double),
                ('sr_mean', ctypes.c_void_p), ('sr_std', ctypes.c_void_p), ('tbias', ctypes.c_void_p),
                ('ebias', ctypes.c_void_p), ('embed', ctypes.c_void_p), ('fit', ctypes.c_void_p)]


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ['g++', '-O3', '-ffast-math', '-march=x86-64-v3', '-fopenmp', '-shared', '-fPIC', '-o', LIB + '.tmp', SRC]
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError('%s is not built (oracle.dp_cpu.build())' % LIB)
        L = ctypes.CDLL(LIB)
        for name in ('dpc_energy_force_f64', 'dpc_energy_force_f32'):
            f = getattr(L, name)
            f.restype = ctypes.c_int32
            f.argtypes = [ctypes.POINTER(_Model), ctypes.c_int32] + [ctypes.c_void_p] * 9
        L.dpc_threads.restype = ctypes.c_int32
        _lib = L
    return _lib


def threads():
    return int(lib().dpc_threads())


class Packed:
    """The model's parameters in the layout dp_cpu.cpp reads (kept alive with the struct)."""

    def __init__(self, params, variables):
        P = variables['params']
        nt = int(params['ntypes'])
        ew = [int(w) for w in params['embed_widths']]
        fw = [int(w) for w in params['fit_widths']]
        emb = []
        for k in range(nt * nt):
            net = P['embedding_net_%d' % k]
            for li in range(len(ew)):
                emb += [np.asarray(net['Dense_%d' % li]['kernel'], np.float64).ravel(),
                        np.asarray(net['Dense_%d' % li]['bias'], np.float64).ravel()]
        fit = []
        for t in range(nt):
            net = P['fitting_net_%d' % t]
            for li in range(len(fw)):
                fit += [np.asarray(net['Dense_%d' % li]['kernel'], np.float64).ravel(),
                        np.asarray(net['Dense_%d' % li]['bias'], np.float64).ravel()]
                if li > 0 and fw[li] == fw[li - 1]:
                    fit.append(np.asarray(net['dt%d' % li], np.float64).ravel())
            last = net['Dense_%d' % len(fw)]
            fit += [np.asarray(last['kernel'], np.float64).ravel(),
                    np.asarray(last.get('bias', np.zeros(1)), np.float64).ravel()]
        self.arrays = {
            'sr_mean': np.ascontiguousarray(params['sr_mean'], np.float64),
            'sr_std': np.ascontiguousarray(params['sr_std'], np.float64),
            'tbias': np.ascontiguousarray(P['Tbias'], np.float64),
            'ebias': np.ascontiguousarray(np.asarray(params['Ebias'], np.float64).reshape(-1)),
            'embed': np.ascontiguousarray(np.concatenate(emb)),
            'fit': np.ascontiguousarray(np.concatenate(fit)),
        }
        m = _Model()
        m.ntypes, m.axis, m.n_embed, m.n_fit = nt, int(params['axis']), len(ew), len(fw)
        for i, w in enumerate(ew):
            m.embed_w[i] = w
        for i, w in enumerate(fw):
            m.fit_w[i] = w
        m.rcut, m.nnbrs = float(params['rcut']), float(params['Nnbrs'])
        m.out_norm = float(params.get('out_norm', 1.0))
        for k, a in self.arrays.items():
            setattr(m, k, a.ctypes.data)
        self.model = m


def neighbour_rows(nbrs_nm, type_idx):
    """(N, sum K_j) sorted atom indices (-1 padding) and K_j from nbrs_nm[i][j] of typed_nlist."""
    tc = np.bincount(np.asarray(type_idx))
    nt = len(tc)
    off = np.concatenate([[0], np.cumsum(tc)])
    ks = [int(np.asarray(nbrs_nm[0][j]).shape[1]) for j in range(nt)]
    rows = []
    for i in range(nt):
        blocks = []
        for j in range(nt):
            b = np.asarray(nbrs_nm[i][j]).astype(np.int64)
            blocks.append(np.where(b < tc[j], b + off[j], -1))
        rows.append(np.concatenate(blocks, 1))
    return np.ascontiguousarray(np.concatenate(rows, 0), np.int32), np.asarray(ks, np.int32)


def energy_force(packed, coord, box, type_idx, nbr_rows, kslots, dtype='f64'):
    """(E, F (N,3) caller order, virial (3,3), per-atom energies caller order)."""
    type_idx = np.asarray(type_idx)
    perm = np.argsort(type_idx, kind='stable')
    N = len(type_idx)
    c = np.ascontiguousarray(np.asarray(coord, np.float64)[perm])
    ts = np.ascontiguousarray(type_idx[perm], np.int32)
    b = np.ascontiguousarray(np.asarray(box, np.float64).reshape(3))
    e = np.zeros(1)
    ea = np.zeros(N)
    f = np.zeros((N, 3))
    w = np.zeros(9)
    fn = lib().dpc_energy_force_f64 if dtype == 'f64' else lib().dpc_energy_force_f32
    rc = fn(ctypes.byref(packed.model), N, ts.ctypes.data, c.ctypes.data, b.ctypes.data, nbr_rows.ctypes.data,
            np.ascontiguousarray(kslots, np.int32).ctypes.data, e.ctypes.data, ea.ctypes.data, f.ctypes.data,
            w.ctypes.data)
    if rc != 0:
        raise RuntimeError('dp_cpu failed: %d' % rc)
    F = np.empty_like(f)
    F[perm] = f
    EA = np.empty_like(ea)
    EA[perm] = ea
    return float(e[0]), F, w.reshape(3, 3), EA
