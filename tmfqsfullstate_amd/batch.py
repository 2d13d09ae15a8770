"""Batched circuits: ctypes binding of the tmfqs_hip_batch_* entry points of
include/tmfqs_hip.h, and the text circuit-list format the batched workload is
committed in.

Circuit-list format: '#' comment lines; 'C <qubits> <gates>' opens a circuit;
then one gate per line: 'H q', 'X q', 'CX c t', 'CP c t theta', 'PH q theta',
'SW a b' (qubit numbering as everywhere in the API: qubit 0 is the most
significant bit of the basis index).
"""

import ctypes
import hashlib
import json
import os

import numpy as np

from . import (C128, Config, DiagTerm, Op, OP_CPHASE, OP_CX, OP_H, OP_PHASE, OP_SWAP, OP_X, check, default_config, lib,
               make_terms)

_KINDS = {"H": OP_H, "X": OP_X, "CX": OP_CX, "CP": OP_CPHASE, "PH": OP_PHASE, "SW": OP_SWAP}


class Circuit(ctypes.Structure):
    _fields_ = [
        ("qubits", ctypes.c_uint),
        ("reserved", ctypes.c_uint32),
        ("init_state", ctypes.c_uint64),
        ("ops", ctypes.POINTER(Op)),
        ("count", ctypes.c_size_t),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("circuits", ctypes.c_size_t),
        ("rounds", ctypes.c_size_t),
        ("passes", ctypes.c_size_t),
        ("items", ctypes.c_size_t),
        ("tile_qubits", ctypes.c_int),
        ("graph", ctypes.c_int),
        ("state_bytes", ctypes.c_double),
    ]


def load_circuit_list(path):
    """Parses a circuit-list file into [(qubits, ops_array, count), ...]."""
    out = []
    cur = None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            t = line.split()
            if t[0] == "C":
                cur = [int(t[1]), int(t[2]), []]
                out.append(cur)
                continue
            kind = _KINDS[t[0]]
            if t[0] == "CP":
                cur[2].append((kind, int(t[1]), int(t[2]), float(t[3])))
            elif t[0] == "PH":
                cur[2].append((kind, int(t[1]), 0, float(t[2])))
            elif t[0] in ("CX", "SW"):
                cur[2].append((kind, int(t[1]), int(t[2])))
            else:
                cur[2].append((kind, int(t[1])))
    circuits = []
    for n, count, seq in out:
        if count != len(seq):
            raise ValueError(f"circuit header says {count} gates, found {len(seq)}")
        arr = (Op * max(1, len(seq)))()
        for i, g in enumerate(seq):
            arr[i].kind = g[0]
            arr[i].q0 = g[1]
            arr[i].q1 = g[2] if len(g) > 2 else 0
            arr[i].theta = g[3] if len(g) > 3 else 0.0
        circuits.append((n, arr, len(seq)))
    return circuits


def tune_key(circuits, precision):
    """Key of a tuned tile width: the precision and every circuit's qubit count
    and gate-kind counts (the shape the tile choice depends on, not the
    angles or the qubits drawn)."""
    h = hashlib.sha1(f"precision={int(precision)}".encode())
    for n, ops, count in circuits:
        kinds = {}
        for i in range(count):
            kinds[ops[i].kind] = kinds.get(ops[i].kind, 0) + 1
        h.update(f";{n}:{sorted(kinds.items())}".encode())
    return h.hexdigest()


def tune_cache_lookup_entry(path, key):
    """The entry stored under `key` in the JSON file `path`, or None."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def tune_cache_lookup(path, key):
    """Tile width stored under `key` in the JSON file `path`, or None."""
    entry = tune_cache_lookup_entry(path, key)
    return None if entry is None else int(entry["tile_qubits"])


def tune_cache_store(path, key, tile_qubits, **info):
    """Adds or replaces `key` in the JSON file `path` (written to a temporary
    file and renamed, so a reader never sees half a file)."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = dict(tile_qubits=int(tile_qubits), **info)
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _bind():
    L = lib()
    if not getattr(L, "_batch_bound", False):
        vp = ctypes.c_void_p
        L.tmfqs_hip_batch_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Circuit), ctypes.c_size_t, ctypes.POINTER(vp)]
        L.tmfqs_hip_batch_tune.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.tmfqs_hip_batch_run.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.tmfqs_hip_batch_export.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
        L.tmfqs_hip_batch_info.argtypes = [vp, ctypes.POINTER(Stats)]
        L.tmfqs_hip_batch_destroy.argtypes = [vp]
        L.tmfqs_hip_batch_adjoint.argtypes = [vp, ctypes.POINTER(DiagTerm), ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
        L._batch_bound = True
    return L


class Batch:
    """Many circuits run as one graph launch (tmfqs_hip_batch).

    tune_cache: path of a JSON file of tuned tile widths.  When it holds one
    for this workload (see tune_key) the batch is planned with it and tune()
    returns it without timing anything; otherwise tune() times the candidates
    and stores the winner there."""

    def __init__(self, circuits, precision=C128, init_states=None, tune_cache=None, **cfg):
        L = _bind()
        self._cache = tune_cache
        self._key = tune_key(circuits, precision) if tune_cache else None
        self.cached_tile = tune_cache_lookup(tune_cache, self._key) if tune_cache else None
        if self.cached_tile is not None and not cfg.get("tile_qubits"):
            cfg["tile_qubits"] = self.cached_tile
        self.n = [c[0] for c in circuits]
        self.nparams = [sum(1 for i in range(c[2]) if c[1][i].kind in (OP_CPHASE, OP_PHASE)) for c in circuits]
        arr = (Circuit * max(1, len(circuits)))()
        for i, (n, ops, count) in enumerate(circuits):
            arr[i].qubits = n
            arr[i].init_state = 0 if init_states is None else int(init_states[i])
            arr[i].ops = ctypes.cast(ops, ctypes.POINTER(Op))
            arr[i].count = count
        c = default_config(precision=precision, **cfg)
        self._h = ctypes.c_void_p()
        check(L.tmfqs_hip_batch_create(ctypes.byref(c), arr, len(circuits), ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            _bind().tmfqs_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def tune(self):
        t = ctypes.c_int()
        check(_bind().tmfqs_hip_batch_tune(self._h, ctypes.byref(t)))
        if self._cache and self.cached_tile is None:
            tune_cache_store(self._cache, self._key, t.value, circuits=len(self.n), qubits=[min(self.n), max(self.n)])
            self.cached_tile = t.value
        return t.value

    def run(self):
        s = ctypes.c_double()
        check(_bind().tmfqs_hip_batch_run(self._h, ctypes.byref(s)))
        return s.value

    def export(self, i):
        out = np.empty(2 << self.n[i], dtype=np.float64)
        check(_bind().tmfqs_hip_batch_export(self._h, i, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def adjoint(self, terms):
        """(expectation per circuit, list of gradient arrays per circuit)."""
        arr, cnt = make_terms(terms)
        e = np.zeros(len(self.n))
        g = np.zeros(max(1, sum(self.nparams)))
        dp = ctypes.POINTER(ctypes.c_double)
        check(_bind().tmfqs_hip_batch_adjoint(self._h, arr, cnt, e.ctypes.data_as(dp), g.ctypes.data_as(dp),
                                              sum(self.nparams)))
        out, off = [], 0
        for k in self.nparams:
            out.append(g[off:off + k].copy())
            off += k
        return e, out

    def info(self):
        s = Stats()
        check(_bind().tmfqs_hip_batch_info(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}
