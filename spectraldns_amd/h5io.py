"""Checkpoint / results files and the kill-file stop (spectralDNS/h5io/HDF5File.py:57-120).

h5py is not importable in this environment, so the files are NumPy ``.npz``
archives under the reference's dataset names:

* checkpoint ``<name>_c.npz``: the spectral checkpoint data under
  ``U/3D/<level>`` with ``tstep`` and ``t``.  It is written to a temporary
  file in the same folder and moved over the old one (``os.replace``), so a
  run killed while writing leaves the previous checkpoint intact.
* results: one archive per written time step, ``<name>_w_<tstep>.npz``, with
  the physical fields under ``<NAME>/3D/<tstep>`` (the reference appends these
  datasets to one HDF5 file); ``read_results`` merges them.  Nothing is kept
  in host memory between writes.

The kill file ``killspectraldns`` in the working directory makes ``update``
checkpoint and exit with status 1 (HDF5File.py:82-92, 107-120).
"""
import glob
import os
import re
import sys

import numpy as np

__all__ = ["HDF5File", "read_checkpoint", "read_results"]


def read_checkpoint(filename):
    """Arrays, t and tstep of a checkpoint written by HDF5File.update: the file
    ``<filename>_c.npz`` (or ``filename`` itself when it ends in .npz) as a
    dict keyed like the reference's HDF5 datasets (``U/3D/0``), with ``t``
    (float) and ``tstep`` (int).  Loads nothing executable (no pickles)."""
    fn = filename if filename.endswith(".npz") else filename + "_c.npz"
    with np.load(fn, allow_pickle=False) as f:
        out = {k: f[k] for k in f.files}
    out["t"] = float(out["t"])
    out["tstep"] = int(out["tstep"])
    return out


def read_results(filename):
    """Every results dataset written under ``filename`` (``<filename>_w_<tstep>.npz``),
    merged into one dict keyed ``<NAME>/3D/<tstep>`` as in the reference's
    results file."""
    out = {}
    pat = re.compile(re.escape(os.path.basename(filename)) + r"_w_(\d+)\.npz$")
    for fn in sorted(glob.glob(glob.escape(filename) + "_w_*.npz")):
        if pat.search(os.path.basename(fn)):
            with np.load(fn, allow_pickle=False) as f:
                out.update({k: f[k] for k in f.files})
    return out


def _host(a):
    try:
        return a.detach().cpu().numpy()
    except AttributeError:
        return np.asarray(a)


def _save_atomic(path, data):
    """np.savez into a temporary file beside `path`, then os.replace onto it."""
    tmp = "%s.tmp%d" % (path, os.getpid())
    try:
        with open(tmp, "wb") as f:
            np.savez(f, **data)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


class HDF5File:
    def __init__(self, filename, checkpoint=None, results=None):
        self.filename = filename
        self.checkpoint = checkpoint or {}
        self.results = results or {}

    def update(self, params, **kw):
        if params.tstep % params.write_result == 0 and self.results.get("data"):
            self.update_components(**kw)
            data = {}
            for name, val in self.results["data"].items():
                arr = val[0] if isinstance(val, (list, tuple)) else val
                data["%s/3D/%d" % (name, params.tstep)] = _host(arr)
            _save_atomic("%s_w_%d.npz" % (self.filename, params.tstep), data)
        kill = self.check_if_kill()
        if params.tstep % params.checkpoint == 0 or kill:
            data = {}
            for level, fields in self.checkpoint.get("data", {}).items():
                for name, arrs in fields.items():
                    data["%s/3D/%s" % (name, level)] = _host(arrs[0])
            data["tstep"] = np.array(params.tstep)
            data["t"] = np.array(float(params.t))
            _save_atomic(self.filename + "_c.npz", data)
            if kill:
                sys.exit(1)

    def update_components(self, **kw):
        pass

    def open(self):
        pass

    def close(self):
        pass

    @staticmethod
    def check_if_kill():
        """True (and the file removed) if 'killspectraldns' is in the working directory."""
        if "killspectraldns" in os.listdir(os.getcwd()):
            os.remove("killspectraldns")
            print("killspectraldns Found! Stopping simulations cleanly by checkpointing...")
            return True
        return False
