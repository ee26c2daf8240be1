"""Bounded run files: a per-file byte cap, atomic JSON and npz writes, and
temporal arrays split over capped NPY shards.

Only the on-disk format is shared with the reference (python/surface_storage.py),
so that either engine reads the other's files:

* a sharded array is a folder with ``array.json`` -- keys ``shape`` (list),
  ``dtype`` (numpy dtype string), ``frames_per_file``, ``max_file_bytes`` --
  and NPY files ``part_000000.npy``, ``part_000001.npy``, ... holding
  ``frames_per_file`` consecutive frames each (the last one the remainder);
* every file, its NPY header included, stays within ``max_file_bytes``; a
  shard's frame count leaves ``HEADER_BYTES`` for the header;
* caps are decimal bytes in (4096, 5e9], default 4e9.
"""
import io
import json
import os
import tempfile
from pathlib import Path

import numpy as np

DEFAULT_FILE_BYTES = 4_000_000_000
HARD_FILE_BYTES = 5_000_000_000
HEADER_BYTES = 4096


def file_limit(value=DEFAULT_FILE_BYTES):
    """Validated per-file cap in bytes."""
    cap = int(value)
    if cap <= HEADER_BYTES or cap > HARD_FILE_BYTES:
        raise ValueError(f'max_file_bytes={cap} outside ({HEADER_BYTES}, {HARD_FILE_BYTES}] decimal bytes')
    return cap


def _replace_with(path, payload):
    """Write ``payload`` (bytes) next to ``path`` and rename it into place."""
    path = Path(path)
    fd, name = tempfile.mkstemp(prefix=path.name + '.', suffix='.partial', dir=path.parent)
    try:
        with os.fdopen(fd, 'wb') as f:
            f.write(payload)
        os.replace(name, path)
    except BaseException:
        if os.path.exists(name):
            os.unlink(name)
        raise


def atomic_json(path, value):
    """Indented JSON, renamed into place so readers never see a partial file."""
    _replace_with(path, (json.dumps(value, indent=2) + '\n').encode())


def capped_npz(path, max_file_bytes=DEFAULT_FILE_BYTES, **arrays):
    """Atomic ``np.savez`` that refuses a file above the cap.  The archive is
    serialised in memory first, so its exact size is known before anything is
    written."""
    cap = file_limit(max_file_bytes)
    buffer = io.BytesIO()
    np.savez(buffer, **arrays)
    size = buffer.getbuffer().nbytes
    if size > cap:
        raise ValueError(f'{Path(path).name}: {size} bytes would be above file cap {cap}')
    _replace_with(path, buffer.getvalue())


class ShardedArray:
    """A (frames, ...) array stored as capped NPY shards plus a manifest.

    ``mode``: ``'w+'`` creates a new array (the folder must hold no manifest
    yet), ``'r+'`` opens one for reading and writing, ``'r'`` read-only.  One
    shard is memory-mapped at a time; shards are created on first write."""

    def __init__(self, folder, shape=None, dtype=np.complex128, mode='r', max_file_bytes=DEFAULT_FILE_BYTES):
        if mode not in ('r', 'r+', 'w+'):
            raise ValueError(f'unknown mode {mode!r}')
        self.folder = Path(folder)
        self.mode = mode
        self._open_part = None
        self._view = None
        manifest = self.folder / 'array.json'
        if mode == 'w+':
            if manifest.exists():
                raise FileExistsError(manifest)
            self.shape = tuple(int(s) for s in shape)
            self.dtype = np.dtype(dtype)
            self.limit = file_limit(max_file_bytes)
            frame_bytes = self.dtype.itemsize * int(np.prod(self.shape[1:], dtype=np.int64))
            self.frames_per_file = (self.limit - HEADER_BYTES) // frame_bytes
            if self.frames_per_file == 0:
                raise ValueError(f'a frame of {frame_bytes} bytes exceeds file cap {self.limit}')
            self.folder.mkdir(parents=True, exist_ok=True)
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

    def _locate(self, index):
        i = int(index)
        if i < 0:
            i += len(self)
        if i < 0 or i >= len(self):
            raise IndexError(f'frame {index} of {len(self)}')
        return divmod(i, self.frames_per_file)

    def _shard(self, part):
        if part == self._open_part:
            return self._view
        self.flush()
        path = self.folder / f'part_{part:06d}.npy'
        if path.exists():
            view = np.load(path, mmap_mode='r' if self.mode == 'r' else 'r+')
        elif self.mode == 'r':
            raise FileNotFoundError(path)
        else:
            rows = min(self.frames_per_file, len(self) - part * self.frames_per_file)
            view = np.lib.format.open_memmap(path, mode='w+', dtype=self.dtype, shape=(rows, *self.shape[1:]))
            if path.stat().st_size > self.limit:
                raise RuntimeError(f'{path.name} would be {path.stat().st_size} bytes, above file cap {self.limit}')
        self._open_part, self._view = part, view
        return view

    def __getitem__(self, index):
        if isinstance(index, slice):
            index = range(*index.indices(len(self)))
        elif np.isscalar(index):
            part, row = self._locate(index)
            return self._shard(part)[row]
        return np.stack([self[int(i)] for i in index]) if len(index) else np.empty((0, *self.shape[1:]), self.dtype)

    def __setitem__(self, index, value):
        if self.mode == 'r':
            raise ValueError('read-only array: opened with mode r')
        part, row = self._locate(index)
        self._shard(part)[row] = value

    def flush(self):
        if self._view is not None and self.mode != 'r':
            self._view.flush()

    def close(self):
        self.flush()
        self._open_part, self._view = None, None
