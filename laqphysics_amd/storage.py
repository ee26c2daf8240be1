"""Bounded run files: per-file byte caps, atomic JSON and npz writes, and
sharded temporal arrays.

Mirror of python/surface_storage.py: ``file_limit`` (:13-16), ``atomic_json``
(:9-11), ``capped_npz`` (:18-26) and ``ShardedArray`` (:28-74), whose on-disk
layout is kept so that either engine can read the other's files: a folder with
``array.json`` ({"shape", "dtype" (numpy dtype string), "frames_per_file",
"max_file_bytes"}) and NPY shards ``part_000000.npy``, ... holding
``frames_per_file`` frames each (the last one the remainder), every shard,
header included, at most ``max_file_bytes`` long.
"""
import json
import os
from pathlib import Path

import numpy as np

DEFAULT_FILE_BYTES = 4_000_000_000
HARD_FILE_BYTES = 5_000_000_000


def file_limit(value=DEFAULT_FILE_BYTES):
    """Per-file byte cap (python/surface_storage.py:13-16)."""
    value = int(value)
    if not 4096 < value <= HARD_FILE_BYTES:
        raise ValueError('max_file_bytes must be >4096 and <=5,000,000,000 (decimal bytes)')
    return value


def atomic_json(path, value):
    path = Path(path)
    temporary = path.with_name(path.name + '.tmp')
    temporary.write_text(json.dumps(value, indent=2) + '\n')
    os.replace(temporary, path)


def capped_npz(path, max_file_bytes=DEFAULT_FILE_BYTES, **arrays):
    """Size-checked atomic npz write (python/surface_storage.py:18-26)."""
    path = Path(path)
    limit = file_limit(max_file_bytes)
    estimate = sum(np.asarray(v).nbytes + 1024 for v in arrays.values()) + 4096
    if estimate > limit:
        raise ValueError(f'{path.name}: checkpoint needs about {estimate} bytes, above file cap {limit}')
    temporary = path.with_name(path.stem + '.tmp.npz')
    np.savez(temporary, **arrays)
    if temporary.stat().st_size > limit:
        temporary.unlink()
        raise ValueError('NPZ exceeds configured file cap')
    os.replace(temporary, path)


class ShardedArray:
    """A (frames, ...) array split over capped NPY shards (python/surface_storage.py:28-74).

    ``mode``: 'w+' creates the folder's manifest (refused if one exists), 'r+'
    and 'r' open an existing one.  Frames are read and written one at a time
    (an integer index, a slice or an index list for reads); one shard is
    memory-mapped at a time."""

    def __init__(self, folder, shape=None, dtype=np.complex128, mode='r', max_file_bytes=DEFAULT_FILE_BYTES):
        if mode not in ('r', 'r+', 'w+'):
            raise ValueError('mode must be r, r+ or w+')
        self.folder = Path(folder)
        self.mode = mode
        self._part = None
        self._map = None
        manifest = self.folder / 'array.json'
        if mode == 'w+':
            if manifest.exists():
                raise FileExistsError(manifest)
            self.folder.mkdir(parents=True, exist_ok=True)
            self.shape = tuple(int(v) for v in shape)
            self.dtype = np.dtype(dtype)
            self.limit = file_limit(max_file_bytes)
            frame = int(np.prod(self.shape[1:])) * self.dtype.itemsize
            # 4096 bytes are left for the NPY header
            self.frames_per_file = (self.limit - 4096) // frame
            if self.frames_per_file < 1:
                raise ValueError('one array frame exceeds file cap; refine storage layout')
            atomic_json(manifest, {'shape': list(self.shape), 'dtype': self.dtype.str,
                                   'frames_per_file': self.frames_per_file, 'max_file_bytes': self.limit})
        else:
            meta = json.loads(manifest.read_text())
            self.shape = tuple(meta['shape'])
            self.dtype = np.dtype(meta['dtype'])
            self.frames_per_file = int(meta['frames_per_file'])
            self.limit = int(meta['max_file_bytes'])

    def __len__(self):
        return self.shape[0]

    def _row(self, index):
        index = int(index)
        if index < 0:
            index += len(self)
        if not 0 <= index < len(self):
            raise IndexError(index)
        part, row = divmod(index, self.frames_per_file)
        if part != self._part:
            self.flush()
            self._map = None
            path = self.folder / f'part_{part:06d}.npy'
            if path.exists():
                self._map = np.load(path, mmap_mode='r' if self.mode == 'r' else 'r+')
            elif self.mode == 'r':
                raise FileNotFoundError(path)
            else:
                frames = min(self.frames_per_file, len(self) - part * self.frames_per_file)
                self._map = np.lib.format.open_memmap(path, mode='w+', dtype=self.dtype,
                                                      shape=(frames, *self.shape[1:]))
                if path.stat().st_size > self.limit:
                    raise RuntimeError('shard exceeded file cap')
            self._part = part
        return row

    def __getitem__(self, index):
        if isinstance(index, slice):
            return np.array([self[i] for i in range(*index.indices(len(self)))])
        if not np.isscalar(index):
            return np.array([self[int(i)] for i in index])
        row = self._row(index)          # opens the shard first
        return self._map[row]

    def __setitem__(self, index, value):
        if self.mode == 'r':
            raise ValueError('read-only array')
        row = self._row(index)
        self._map[row] = value

    def flush(self):
        if self._map is not None and self.mode != 'r':
            self._map.flush()

    def close(self):
        self.flush()
        self._map = None
        self._part = None
