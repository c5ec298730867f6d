"""Streaming XYZ trajectory dumps for `Simulation.run(dump_prefix=...)`.

The argument rules, file naming and the extended-XYZ frame text follow
deepmd_jax/md_utils.py:20-256 (prepare_dumps, write_dump_frame, _write_xyz_frame):
`{prefix}_{kind}.xyz` per dump kind ('position', 'velocity', 'centroid'), one file per
bead for positions and velocities under PIMD (`{prefix}_{kind}_bead{b}.xyz`), frames
every `dump_interval` steps (default: the report interval) with the step number in the
comment line, existing files truncated unless dump_mode='append'.

Frames are written as the reference writes them with jax's default 32-bit arrays: the
values rounded to float32 and printed with 8 decimals (15 for float64 frames).  The text
is formatted and appended by a writer thread while the device computes the following
steps; the queue is bounded so a writer that falls behind holds the run back instead of
growing without limit.
"""
import glob
import os
import queue
import threading

import numpy as np

DUMP_KINDS = ('position', 'velocity', 'centroid')
DUMP_MODES = ('overwrite', 'append')

# Chemical symbols by atomic number (index 0 is the placeholder 'X'), for models that store
# `chemical_types` (deepmd_jax/md.py:186-187 takes them from ase.data.chemical_symbols).
CHEMICAL_SYMBOLS = (
    'X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr '
    'Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb '
    'Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf '
    'Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og').split()


def normalize_dump_mode(dump_mode):
    if dump_mode not in DUMP_MODES:
        raise ValueError("dump_mode must be 'overwrite' or 'append'; got %r." % (dump_mode,))
    return dump_mode


def normalize_dump_content(dump_content, n_bead):
    """Expand 'all' / 'positions' and validate the kinds (deepmd_jax/md_utils.py:32-68)."""
    is_pimd = n_bead > 1
    if isinstance(dump_content, str):
        items = (dump_content,)
    else:
        try:
            items = tuple(dump_content)
        except TypeError as exc:
            raise ValueError("dump_content must be 'all', 'positions', a dump kind string, "
                             "or an iterable of dump kind strings.") from exc
    if not items:
        raise ValueError('dump_content must not be empty when dump_prefix is set.')
    bad = [it for it in items if not isinstance(it, str)]
    if bad:
        raise ValueError('dump_content entries must be strings; got %r.' % (bad,))
    kinds = []
    for it in items:
        if it == 'all':
            kinds += list(DUMP_KINDS) if is_pimd else ['position', 'velocity']
        elif it == 'positions':
            kinds += ['position', 'centroid'] if is_pimd else ['position']
        else:
            kinds.append(it)
    unknown = sorted(set(kinds) - set(DUMP_KINDS))
    if unknown:
        raise ValueError("Unknown dump_content entries %s. Allowed entries are %s, 'positions', or 'all'."
                         % (unknown, list(DUMP_KINDS)))
    if 'centroid' in kinds and not is_pimd:
        raise ValueError("dump_content includes 'centroid', which is only meaningful when n_bead > 1.")
    return tuple(k for k in DUMP_KINDS if k in kinds)


def requested_paths(dump_prefix, dump_content, n_bead):
    if dump_prefix is None:
        if not (isinstance(dump_content, str) and dump_content in ('all', 'positions')):
            raise ValueError('dump_content requires dump_prefix to be set.')
        return {}
    prefix = os.fspath(dump_prefix)
    if prefix == '':
        raise ValueError('dump_prefix must be a non-empty path prefix.')
    return {k: '%s_%s.xyz' % (prefix, k) for k in normalize_dump_content(dump_content, n_bead)}


def bead_path(path, bead):
    root, ext = os.path.splitext(os.fspath(path))
    return '%s_bead%d%s' % (root, bead, ext)


def _remove_target(path, bead_split):
    if os.path.exists(path):
        os.remove(path)
    if bead_split:
        root, ext = os.path.splitext(path)
        for p in glob.glob('%s_bead*%s' % (root, ext)):
            if os.path.exists(p):
                os.remove(p)


def prepare_dumps(type_idx, type_symbols, n_bead, default_interval, dump_interval=None, dump_prefix=None,
                  dump_content='positions', dump_mode='overwrite'):
    """Validate the dump arguments; returns ([(kind, path, bead or None)], symbols, interval)
    or ([], None, None) when nothing is requested (deepmd_jax/md_utils.py:182-232)."""
    req = requested_paths(dump_prefix, dump_content, n_bead)
    if not req:
        return [], None, None
    dump_mode = normalize_dump_mode(dump_mode)
    if type_symbols is None:
        raise ValueError('Simulation.run(dump_prefix=...) requires `type_symbols=[...]` at Simulation(...) '
                         'construction so XYZ output can label atoms.')
    is_pimd = n_bead > 1
    interval = int(dump_interval) if dump_interval is not None else int(default_interval)
    if interval <= 0:
        raise ValueError('dump_interval must be positive; got %d' % interval)
    symbols = [str(type_symbols[t]) for t in np.asarray(type_idx)]
    dumps = []
    for kind, path in req.items():
        split = is_pimd and kind != 'centroid'
        if dump_mode == 'overwrite':
            _remove_target(path, split)
        if split:
            dumps += [(kind, bead_path(path, b), b) for b in range(n_bead)]
        else:
            dumps.append((kind, path, None))
    return dumps, symbols, interval


def target_label(dumps, dump_prefix):
    if dump_prefix is not None:
        return '%s_*' % (dump_prefix,)
    return ', '.join(d[1] for d in dumps)


def format_frame(symbols, data, box, step=None):
    """Extended-XYZ text of one frame (deepmd_jax/md_utils.py:115-136): atom count, the comment
    line (Lattice of the cell row by row, Properties, step, pbc), then one
    `%-2s x y z` line per atom.  data (N,3) float32 or float64."""
    data = np.asarray(data)
    fmt = '.8f' if data.dtype.itemsize <= 4 else '.15f'
    box = np.asarray(box, np.float64)
    cell = np.diag(box) if box.shape == (3,) else box.reshape(3, 3)
    comment = []
    if cell.any():
        comment.append('Lattice="%s"' % ' '.join(format(float(v), fmt) for v in cell.reshape(-1)))
    comment.append('Properties=species:S:1:pos:R:3')
    if step is not None:
        comment.append('step=%d' % int(step))
    comment.append('pbc="T T T"')
    line = '%-2s %' + fmt + ' %' + fmt + ' %' + fmt + '\n'
    rows = data.tolist()
    body = ''.join([line % (s, r[0], r[1], r[2]) for s, r in zip(symbols, rows)])
    return '%d\n%s\n%s' % (len(symbols), ' '.join(comment), body)


class DumpWriter:
    """Appends frames to the dump files from a writer thread.  put() takes arrays the caller
    no longer modifies; close() waits for the queue and re-raises a writer error."""

    def __init__(self, dumps, symbols, dtype=np.float32, depth=16):
        self.dumps, self.symbols, self.dtype = dumps, symbols, dtype
        self.q = queue.Queue(maxsize=depth)
        self.error = None
        self.thread = threading.Thread(target=self._loop, daemon=True)
        self.thread.start()

    def _loop(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            if self.error is not None:
                continue
            try:
                self._write(*item)
            except BaseException as exc:  # reported by close()
                self.error = exc

    def _write(self, pos, vel, cen, box, step):
        src = {'position': pos, 'velocity': vel, 'centroid': cen}
        for kind, path, bead in self.dumps:
            d = src[kind]
            if d is None:
                raise ValueError('Missing %s frame for XYZ dump.' % kind)
            d = np.asarray(d)
            if bead is not None:
                d = d[bead]
            elif d.ndim == 3:
                d = d.reshape(-1, 3)
            with open(path, 'a', encoding='utf-8') as fh:
                fh.write(format_frame(self.symbols, d.astype(self.dtype, copy=False), box.astype(self.dtype), step))

    def put(self, pos, vel, cen, box, step):
        if self.error is not None:
            self.close()
        self.q.put((pos, vel, cen, np.asarray(box), step))

    def close(self):
        if self.thread.is_alive():
            self.q.put(None)
            self.thread.join()
        if self.error is not None:
            err, self.error = self.error, None
            raise err
