"""ctypes binding of the FP64 CPU oracle (oracle/libgrn_oracle.so).

TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
benchmark's CPU-baseline leg use it as the checker.  The product library
never imports it.  The C restatement and its reference file:line map are in
grn_oracle.c / grn_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrn_oracle.so")

GENE_DTYPE = np.dtype({"names": ["label", "init", "lim", "deg", "dif"],
                       "formats": [np.int32, np.float64, np.float64, np.float64, np.float64],
                       "offsets": [0, 8, 16, 24, 32], "itemsize": 40})
LINK_DTYPE = np.dtype({"names": ["regulator", "regulated", "dis_const", "hill_coef", "is_and"],
                       "formats": [np.int32, np.int32, np.float64, np.float64, np.int32],
                       "offsets": [0, 4, 8, 16, 24], "itemsize": 32})
RESULT_DTYPE = np.dtype({"names": ["fitness", "raw_error", "change", "n_products", "steps",
                                   "experiments"],
                         "formats": [np.float64, np.float64, np.float64, np.int32, np.int32,
                                     np.int32],
                         "offsets": [0, 8, 16, 24, 28, 32], "itemsize": 40})
FITNESS_REF, FITNESS_MSE = 0, 1


class _Params(ctypes.Structure):
    _fields_ = [("dim_x", ctypes.c_int), ("dim_y", ctypes.c_int), ("n_in", ctypes.c_int),
                ("n_tgt", ctypes.c_int), ("n_steps", ctypes.c_int), ("dt", ctypes.c_double),
                ("min_conc", ctypes.c_double), ("min_conc_change", ctypes.c_double),
                ("dist_error_threshold", ctypes.c_double), ("error_precision", ctypes.c_int),
                ("fitness_kind", ctypes.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        lib.gro_ceil_s.argtypes = [ctypes.c_double, ctypes.c_int]
        lib.gro_ceil_s.restype = ctypes.c_double
        lib.gro_labels_in_use.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, vp,
                                          ctypes.c_int]
        lib.gro_evaluate.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, vp, vp, ctypes.c_int,
                                     vp, ctypes.c_int, vp, ctypes.c_double, vp, vp]
        lib.gro_evaluate_batch_dt.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, vp, vp,
                                              ctypes.c_int, vp, vp, vp, vp, vp, vp, ctypes.c_int,
                                              vp]
        lib.gro_decode_rgb.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp]
        _lib = lib
    return _lib


def params(dim_x, dim_y, n_in, n_tgt, n_steps, dt, min_conc=0.0, min_conc_change=-1.0,
           dist_error_threshold=0.0, error_precision=0, fitness_kind=FITNESS_MSE) -> _Params:
    return _Params(dim_x, dim_y, n_in, n_tgt, n_steps, dt, min_conc, min_conc_change,
                   dist_error_threshold, error_precision, fitness_kind)


def from_simparams(sp) -> _Params:
    """Builds oracle params from an object with the SimParams field names."""
    return params(sp.dim_x, sp.dim_y, sp.n_input_morphogens, sp.n_target_morphogens,
                  sp.num_sim_steps, sp.dt, sp.min_conc, sp.min_conc_change,
                  sp.dist_error_threshold, sp.error_precision, sp.fitness_kind)


def ceil_s(n: float, d: int) -> float:
    return load().gro_ceil_s(n, d)


def labels_in_use(genes, links, n_morph):
    genes = np.ascontiguousarray(genes, GENE_DTYPE)
    links = np.ascontiguousarray(links, LINK_DTYPE)
    out = np.zeros(256, np.int32)
    n = load().gro_labels_in_use(len(genes), genes.ctypes.data, len(links),
                                 links.ctypes.data if len(links) else None, n_morph,
                                 out.ctypes.data, len(out))
    if n < 0:
        raise ValueError("too many labels in use")
    return out[:n].copy()


def evaluate(p: _Params, init, target, genes, links, max_error=float("inf"), want_fields=False):
    """EvaluatorMorphogens::evaluate for one model -> (result, fields|None)."""
    init = np.ascontiguousarray(init, np.float64)
    target = np.ascontiguousarray(target, np.float64)
    genes = np.ascontiguousarray(genes, GENE_DTYPE)
    links = np.ascontiguousarray(links, LINK_DTYPE)
    n_exp = init.shape[0] if init.size else target.shape[0]
    res = np.zeros(1, RESULT_DTYPE)
    fields = np.zeros(64 * p.dim_x * p.dim_y) if want_fields else None
    rc = load().gro_evaluate(ctypes.byref(p), n_exp, init.ctypes.data if init.size else None,
                             target.ctypes.data if target.size else None, len(genes),
                             genes.ctypes.data, len(links),
                             links.ctypes.data if len(links) else None, max_error,
                             res.ctypes.data, fields.ctypes.data if want_fields else None)
    if rc:
        raise ValueError(f"oracle rejected the model ({rc})")
    r = res[0]
    if want_fields:
        g = int(r["n_products"])
        fields = fields[: g * p.dim_x * p.dim_y].reshape(g, p.dim_y, p.dim_x)
    return r, fields


def evaluate_batch(p: _Params, init, target, models, max_errors=None, dt=None, n_threads=1):
    """Model-parallel evaluation on n_threads host threads."""
    init = np.ascontiguousarray(init, np.float64)
    target = np.ascontiguousarray(target, np.float64)
    gene_off = np.zeros(len(models) + 1, np.int32)
    link_off = np.zeros(len(models) + 1, np.int32)
    for i, m in enumerate(models):
        gene_off[i + 1] = gene_off[i] + len(m.genes)
        link_off[i + 1] = link_off[i] + len(m.links)
    # fill preallocated arrays: np.concatenate would drop the dtype's padding
    genes = np.zeros(max(int(gene_off[-1]), 1), GENE_DTYPE)
    links = np.zeros(max(int(link_off[-1]), 1), LINK_DTYPE)
    for i, m in enumerate(models):
        for name in GENE_DTYPE.names:
            genes[name][gene_off[i]:gene_off[i + 1]] = np.asarray(m.genes)[name]
        for name in LINK_DTYPE.names:
            links[name][link_off[i]:link_off[i + 1]] = np.asarray(m.links)[name]
    me = None if max_errors is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(max_errors, np.float64), (len(models),)))
    dts = None if dt is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(dt, np.float64), (len(models),)))
    out = np.zeros(len(models), RESULT_DTYPE)
    n_exp = init.shape[0] if init.size else target.shape[0]
    rc = load().gro_evaluate_batch_dt(ctypes.byref(p), n_exp, init.ctypes.data if init.size else None,
                                      target.ctypes.data if target.size else None, len(models),
                                      gene_off.ctypes.data, genes.ctypes.data, link_off.ctypes.data,
                                      links.ctypes.data,
                                      None if me is None else me.ctypes.data,
                                      None if dts is None else dts.ctypes.data, n_threads,
                                      out.ctypes.data)
    if rc:
        raise ValueError(f"oracle rejected a model ({rc})")
    return out


def decode_rgb(images, n_planes):
    """SimState::loadMorphologyImage (simstate.cpp:71-90) of each image in
    images[n][H][W] (uint32 0xAARRGGBB, scan lines top to bottom) ->
    planes [n][n_planes][H][W] (float64)."""
    img = np.ascontiguousarray(images, np.uint32)
    n, H, W = img.shape
    out = np.zeros((n, n_planes, H, W), np.float64)
    for e in range(n):
        o = np.zeros((n_planes, H, W), np.float64)
        rc = load().gro_decode_rgb(W, H, img[e].ctypes.data, n_planes, o.ctypes.data)
        if rc:
            raise ValueError("bad image size or plane count")
        out[e] = o
    return out
