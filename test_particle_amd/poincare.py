"""Host side of the Poincare map (the reference's run_type 2).

The reference reads its "Poincare map configuration" block from init.dat
(config_poincare_map, src/poincare_map.c:357-514), lays out the start points
(:83-158), traces them (:120-185) and writes data/poincare_map.h5 (:188-246).
This module reads and writes that block, builds tp_pmap_params and the start
points, and runs the map on a Tracker (tp_gpu_poincare_map); the tracing
itself always runs on the GPU.
"""
import re

from . import abi
from .gpu import pmap_starts

# the block's lines in the order config_poincare_map scans them
_HEAD = [("face", "Which face as the cross section (0: x, 1: y, 2:z):", int),
         ("step", "Field-line tracking step:", float),
         ("length_max", "Maximum tracking length:", float),
         ("ncross", "Number of crossing points:", int),
         ("single", "Whether to get the map only for a single point:", int),
         ("x0", "x-position of the point:", float),
         ("y0", "y-position of the point:", float),
         ("z0", "z-position of the point:", float)]
_PLANE = [("min_h", "min_horizontal:", float), ("max_h", "max_horizontal:", float),
          ("min_v", "min_vertical:", float), ("max_v", "max_vertical:", float),
          ("nh", "nhorizontal:", int), ("nv", "nvertical:", int)]

_NUM = re.compile(r"^\s*([-+0-9.eEinfINFaA]+)")


def _value(line, label, conv):
    head = line.strip()
    if not head.startswith(label):
        raise ValueError(f"expected '{label}', found '{line.rstrip()}'")
    m = _NUM.match(head[len(label):])
    if not m:
        raise ValueError(f"no value after '{label}'")
    return conv(float(m.group(1))) if conv is int else conv(m.group(1))


def read_config(path):
    """The Poincare block of an init.dat as a dict with keys face, step,
    length_max, ncross, single, x0, y0, z0 and, for a plane of start points
    (single = 0), min_h, max_h, min_v, max_v, nh, nv.  Like the reference,
    the block starts after the line containing "Poincare map configuration"
    and the plane section after the line containing "for a 2D plane"."""
    with open(path) as f:
        lines = f.read().splitlines()
    i = next((k for k, ln in enumerate(lines) if "Poincare map configuration" in ln), None)
    if i is None:
        raise ValueError("no Poincare map configuration block")
    cfg = {}
    pos = i + 1
    for key, label, conv in _HEAD:
        cfg[key] = _value(lines[pos], label, conv)
        pos += 1
    if cfg["face"] not in (0, 1, 2):
        raise ValueError("wrong face for cross section")
    if not cfg["single"]:
        j = next((k for k in range(pos, len(lines)) if "for a 2D plane" in lines[k]), None)
        if j is None:
            raise ValueError("no 2D plane section")
        pos = j + 1
        for key, label, conv in _PLANE:
            cfg[key] = _value(lines[pos], label, conv)
            pos += 1
    return cfg


def write_config(path, cfg):
    """Writes the block in init.dat's format (doubles with repr, so that the
    reference's %lf reads back the same values)."""
    r = lambda v: repr(float(v))  # noqa: E731
    out = ["-" * 25 + " Poincare map configuration " + "-" * 25,
           f"Which face as the cross section (0: x, 1: y, 2:z): {int(cfg['face'])}",
           f"Field-line tracking step: {r(cfg['step'])}",
           f"Maximum tracking length: {r(cfg['length_max'])}",
           f"Number of crossing points: {int(cfg['ncross'])}",
           f"Whether to get the map only for a single point: {int(cfg['single'])}",
           f"x-position of the point: {r(cfg['x0'])}",
           f"y-position of the point: {r(cfg['y0'])}",
           f"z-position of the point: {r(cfg['z0'])}",
           "---------- for a 2D plane --------",
           f"min_horizontal: {r(cfg.get('min_h', 0.0))}",
           f"max_horizontal: {r(cfg.get('max_h', 0.0))}",
           f"min_vertical: {r(cfg.get('min_v', 0.0))}",
           f"max_vertical: {r(cfg.get('max_v', 0.0))}",
           f"nhorizontal: {int(cfg.get('nh', 1))}",
           f"nvertical: {int(cfg.get('nv', 1))}",
           "-" * 79]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def params(cfg, lo, hi, mode=abi.TP_PMAP_REF_COMPAT, chunk=0):
    """tp_pmap_params of a configuration; lo/hi the simulation domain."""
    return abi.make_pmap_params(face=cfg["face"], ncross=cfg["ncross"], step=cfg["step"],
                                length_max=cfg["length_max"], lo=lo, hi=hi, mode=mode,
                                chunk=chunk)


def starts(cfg):
    """Start points (npoints, 3) in the reference's order."""
    return pmap_starts(cfg["face"], (cfg["x0"], cfg["y0"], cfg["z0"]),
                       single=bool(cfg["single"]), min_h=cfg.get("min_h", 0.0),
                       max_h=cfg.get("max_h", 0.0), min_v=cfg.get("min_v", 0.0),
                       max_v=cfg.get("max_v", 0.0), nh=cfg.get("nh", 1), nv=cfg.get("nv", 1))


def run(tracker, cfg, lo, hi, mode=abi.TP_PMAP_REF_COMPAT, chunk=0):
    """The map of the tracker's current field: dict(xpos, ypos, zpos,
    ncrossed, nsteps), the position arrays (npoints, ncross + 1) as the
    reference's datasets."""
    return tracker.poincare_map(params(cfg, lo, hi, mode, chunk), starts(cfg))
