"""Autotuned launch geometry, cached per device architecture.

A member's launch geometry (column-tile width, vertical chunking, XCD
ordering) changes which workgroup computes a cell, never the result, so it is
chosen by timing (nimbus_hip_member_autotune) and remembered per
(architecture, nx, ny, nz, precision, acoustic mode) in a JSON file: one file
per architecture, nimbus_amd/tuning/<arch>.json by default. Shapes seen
before are set without timing; new ones are tuned once and added.
"""
import json
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DIR = os.path.join(HERE, "tuning")


class TuningCache:
    def __init__(self, arch, directory=None, trials=2, persist=True):
        self.arch = arch
        self.path = os.path.join(directory or DEFAULT_DIR, f"{arch}.json")
        self.trials = int(trials)
        self.persist = persist
        self.hits = 0
        self.tuned = 0
        self._lock = threading.Lock()
        try:
            with open(self.path) as f:
                self.table = json.load(f)
        except (OSError, ValueError):
            self.table = {}

    @staticmethod
    def key(grid, precision, mode):
        return f"{grid.nx}x{grid.ny}x{grid.nz}:p{int(precision)}:a{int(mode)}"

    def apply(self, member, mode):
        """Sets the member's launch geometry from the cache, tuning it first
        when its shape is new. Returns (tile_w, k_chunks, xcd_remap)."""
        k = self.key(member.grid, member.precision, mode)
        with self._lock:
            e = self.table.get(k)
        if e is None:
            w, c, r, ms = member.autotune(self.trials)
            e = {"tile_w": w, "k_chunks": c, "xcd_remap": r, "ms_per_step": round(ms, 5)}
            with self._lock:
                self.table[k] = e
                self.tuned += 1
            if self.persist:
                self.save()
        else:
            member.set_tiling(e["tile_w"], e["k_chunks"], e["xcd_remap"])
            self.hits += 1
        return e["tile_w"], e["k_chunks"], e["xcd_remap"]

    def save(self):
        with self._lock:
            data = dict(sorted(self.table.items()))
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, self.path)
