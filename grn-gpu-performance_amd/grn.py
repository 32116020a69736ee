"""Python binding of libgrn_hip.so (the C ABI declared in include/grn.h).

Mirrors the reference's GPU evaluator interface: a Context is
CEvaluatorMorphogens (Src/GPUSimulator/cevaluatormorphogens.h:21-62) -- device
init plus experiment upload -- and Context.evaluate(model, max_error) is
CEvaluatorMorphogens::evaluate (cevaluatormorphogens.cpp:233-245).  Models
are the reference's Model (genes = ModelProd, links = ModelLink) as numpy
structured arrays.

The library must exist (build it with ``python -c "import
__graft_entry__ as g; g.build()"`` or ``make -C grn-gpu-performance_amd``);
there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrn_hip.so")

FITNESS_REF = 0
FITNESS_MSE = 1
PATH_AUTO, PATH_LDS, PATH_TILED, PATH_EXACT = 0, 1, 2, 3
LAUNCH_EAGER, LAUNCH_GRAPH = 0, 1
MAX_GENES = 16
MAX_LINKS = 96
NOALLOC_MAX_MODELS = 256

GENE_DTYPE = np.dtype({"names": ["label", "init", "lim", "deg", "dif"],
                       "formats": [np.int32, np.float64, np.float64, np.float64, np.float64],
                       "offsets": [0, 8, 16, 24, 32], "itemsize": 40})
LINK_DTYPE = np.dtype({"names": ["regulator", "regulated", "dis_const", "hill_coef", "is_and"],
                       "formats": [np.int32, np.int32, np.float64, np.float64, np.int32],
                       "offsets": [0, 4, 8, 16, 24], "itemsize": 32})

EXPORTED = [
    "grn_abi_version", "grn_last_error", "grn_device_count", "grn_create", "grn_destroy",
    "grn_add_dataset", "grn_decode_rgb", "grn_add_dataset_rgb", "grn_evaluate", "grn_evaluate_batch",
    "grn_evaluate_batch_noalloc", "grn_stage_batch", "grn_run_staged",
    "grn_staged_destroy", "grn_simulate_fields",
    "grn_get_stats", "grn_test_performance", "grn_set_launch_mode", "grn_autotune", "grn_apply_tuning", "grn_clear_tuning",
    "grn_ceil_s", "grn_model_parse", "grn_model_format", "grn_model_info",
    "grn_copy_probe",
    "grn_search_default_params", "grn_search_create", "grn_search_destroy", "grn_search_pending",
    "grn_search_advance", "grn_search_run_generation", "grn_search_best",
    "grn_search_pareto", "grn_search_pareto_model",
    "grn_search_set_oversample", "grn_search_candidates", "grn_search_screen",
    "grn_search_set_screen_fraction", "grn_search_set_screen_policy", "grn_search_record_raw",
    "grn_search_run_generation_screened", "grn_search_run_generation_coarse",
    "grn_search_run_generation_scaled", "grn_search_rescore",
    "grn_surrogate_default_params", "grn_surrogate_create", "grn_surrogate_destroy",
    "grn_surrogate_n_params", "grn_surrogate_encode", "grn_surrogate_add", "grn_surrogate_train",
    "grn_surrogate_set_pair_weight", "grn_surrogate_predict", "grn_surrogate_predict_features", "grn_surrogate_gradient",
    "grn_surrogate_get_params", "grn_surrogate_set_params", "grn_surrogate_get_stats",
    "grn_mlp_gemm_bf16",
]

MLP_FEATURES = 468
MLP_WIDTH = 512

MIGRATION_PAIRWISE, MIGRATION_RING, MIGRATION_NONE = 0, 1, 2
SCREEN_ABSOLUTE, SCREEN_DELTA, SCREEN_HALF = 0, 1, 2


class GrnError(RuntimeError):
    """A grn_* call returned a negative grn_status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"grn error {code}: {msg}")
        self.code = code


class _SimParams(ctypes.Structure):
    _fields_ = [("dim_x", ctypes.c_int), ("dim_y", ctypes.c_int),
                ("n_input_morphogens", ctypes.c_int), ("n_target_morphogens", ctypes.c_int),
                ("num_sim_steps", ctypes.c_int), ("dt", ctypes.c_double),
                ("min_conc", ctypes.c_double), ("min_conc_change", ctypes.c_double),
                ("dist_error_threshold", ctypes.c_double), ("error_precision", ctypes.c_int),
                ("fitness_kind", ctypes.c_int)]


class _Model(ctypes.Structure):
    _fields_ = [("n_genes", ctypes.c_int), ("genes", ctypes.c_void_p),
                ("n_links", ctypes.c_int), ("links", ctypes.c_void_p)]


class _Result(ctypes.Structure):
    _fields_ = [("fitness", ctypes.c_double), ("raw_error", ctypes.c_double),
                ("change", ctypes.c_double), ("n_products", ctypes.c_int),
                ("steps", ctypes.c_int), ("experiments", ctypes.c_int),
                ("reserved", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("host_ms", ctypes.c_double),
                ("cell_steps", ctypes.c_double), ("alg_flops", ctypes.c_double),
                ("n_launches", ctypes.c_int), ("n_models", ctypes.c_int),
                ("launch_mode", ctypes.c_int), ("graph_replayed", ctypes.c_int)]


class _SearchParams(ctypes.Structure):
    _fields_ = [("n_islands", ctypes.c_int), ("island_size", ctypes.c_int),
                ("n_morphogens", ctypes.c_int), ("n_max_products", ctypes.c_int),
                ("n_max_links", ctypes.c_int), ("init_products_min", ctypes.c_int),
                ("init_products_max", ctypes.c_int), ("migration", ctypes.c_int),
                ("migration_period", ctypes.c_int), ("migration_fraction", ctypes.c_double),
                ("crossover_rate", ctypes.c_double),
                ("lim_lo", ctypes.c_double), ("lim_hi", ctypes.c_double),
                ("deg_lo", ctypes.c_double), ("deg_hi", ctypes.c_double),
                ("dif_prob", ctypes.c_double), ("dif_lo", ctypes.c_double),
                ("dif_hi", ctypes.c_double), ("k_lo", ctypes.c_double), ("k_hi", ctypes.c_double),
                ("n_lo", ctypes.c_double), ("n_hi", ctypes.c_double),
                ("dt_rule", ctypes.c_int), ("dt_scale", ctypes.c_double), ("dt_cap", ctypes.c_double),
                ("seed", ctypes.c_ulonglong), ("n_generations", ctypes.c_int), ("reserved", ctypes.c_int)]


class _GenerationStats(ctypes.Structure):
    _fields_ = [("generation", ctypes.c_int), ("n_evaluated", ctypes.c_int),
                ("min_fit", ctypes.c_double), ("mean_fit", ctypes.c_double),
                ("max_fit", ctypes.c_double), ("best_complexity", ctypes.c_int),
                ("best_island", ctypes.c_int), ("pareto_size", ctypes.c_int),
                ("done", ctypes.c_int), ("last_generation", ctypes.c_int),
                ("best_perfect_complexity", ctypes.c_int)]


class _ParetoEntry(ctypes.Structure):
    _fields_ = [("error", ctypes.c_double), ("complexity", ctypes.c_int), ("generation", ctypes.c_int),
                ("island", ctypes.c_int), ("n_genes", ctypes.c_int), ("n_links", ctypes.c_int),
                ("reserved", ctypes.c_int)]


TUNE_ONCHIP_EXACT, TUNE_ONCHIP_ANY, TUNE_TILED = 0, 1, 2
TUNE_MAX_CANDIDATES = 24


class _TuneCandidate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("cells_per_thread", ctypes.c_int), ("block", ctypes.c_int),
                ("genes_lo", ctypes.c_int), ("genes_hi", ctypes.c_int), ("tile_w", ctypes.c_int),
                ("tile_h", ctypes.c_int), ("tile_sx", ctypes.c_int), ("tile_sy", ctypes.c_int),
                ("halo", ctypes.c_int), ("ms", ctypes.c_double), ("cell_steps_per_s", ctypes.c_double)]


class _TuneClass(ctypes.Structure):
    _fields_ = [("dim_x", ctypes.c_int), ("dim_y", ctypes.c_int), ("genes", ctypes.c_int),
                ("n_target_morphogens", ctypes.c_int), ("guard", ctypes.c_int), ("n_models", ctypes.c_int),
                ("n_timed", ctypes.c_int), ("n_candidates", ctypes.c_int), ("chosen", ctypes.c_int),
                ("reserved", ctypes.c_int), ("cand", _TuneCandidate * TUNE_MAX_CANDIDATES)]


def _tune_to_dict(t: _TuneClass) -> dict:
    d = {k: getattr(t, k) for k, _ in _TuneClass._fields_ if k != "cand"}
    d["cand"] = [{k: getattr(t.cand[j], k) for k, _ in _TuneCandidate._fields_} for j in range(t.n_candidates)]
    return d


def _tune_from_dict(d: dict) -> _TuneClass:
    t = _TuneClass()
    for k, _ in _TuneClass._fields_:
        if k != "cand":
            setattr(t, k, int(d[k]))
    for j, c in enumerate(d["cand"][:TUNE_MAX_CANDIDATES]):
        for k, _ in _TuneCandidate._fields_:
            setattr(t.cand[j], k, c[k])
    return t


class _SurrogateParams(ctypes.Structure):
    _fields_ = [("n_morphogens", ctypes.c_int), ("batch", ctypes.c_int), ("capacity", ctypes.c_int),
                ("reserved", ctypes.c_int), ("lr", ctypes.c_double), ("beta1", ctypes.c_double),
                ("beta2", ctypes.c_double), ("eps", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


class _SurrogateStats(ctypes.Structure):
    _fields_ = [("samples", ctypes.c_long), ("train_steps", ctypes.c_long), ("last_loss", ctypes.c_double),
                ("mu", ctypes.c_double), ("sigma", ctypes.c_double), ("train_ms", ctypes.c_double),
                ("predict_ms", ctypes.c_double)]


class _ScreenStats(ctypes.Structure):
    _fields_ = [("n_candidates", ctypes.c_int), ("n_simulated", ctypes.c_int), ("spearman", ctypes.c_double),
                ("mae", ctypes.c_double), ("loss", ctypes.c_double), ("predict_ms", ctypes.c_double),
                ("train_ms", ctypes.c_double), ("simulate_ms", ctypes.c_double), ("breed_ms", ctypes.c_double),
                ("samples", ctypes.c_long), ("record_ms", ctypes.c_double)]


RESULT_DTYPE = np.dtype({"names": ["fitness", "raw_error", "change", "n_products", "steps",
                                   "experiments", "reserved"],
                         "formats": [np.float64, np.float64, np.float64, np.int32, np.int32,
                                     np.int32, np.int32],
                         "offsets": [0, 8, 16, 24, 28, 32, 36], "itemsize": 40})

_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libgrn_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build the HIP library first")
    lib = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    lib.grn_abi_version.restype = ctypes.c_int
    lib.grn_last_error.restype = ctypes.c_char_p
    lib.grn_device_count.argtypes = [ip]
    lib.grn_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.grn_destroy.argtypes = [vp]
    lib.grn_destroy.restype = None
    lib.grn_add_dataset.argtypes = [vp, ctypes.POINTER(_SimParams), ctypes.c_int, vp, vp, ip]
    lib.grn_decode_rgb.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp]
    lib.grn_add_dataset_rgb.argtypes = [vp, ctypes.POINTER(_SimParams), ctypes.c_int, vp, vp, ip]
    lib.grn_evaluate.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Model), ctypes.c_double,
                                 ctypes.POINTER(_Result)]
    lib.grn_evaluate_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp]
    lib.grn_evaluate_batch_noalloc.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
    lib.grn_stage_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, ctypes.POINTER(vp)]
    lib.grn_run_staged.argtypes = [vp, vp, vp]
    lib.grn_staged_destroy.argtypes = [vp]
    lib.grn_staged_destroy.restype = None
    lib.grn_simulate_fields.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Model), ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, vp, ctypes.c_long, ip]
    lib.grn_get_stats.argtypes = [vp, ctypes.POINTER(_Stats)]
    lib.grn_set_launch_mode.argtypes = [vp, ctypes.c_int]
    lib.grn_test_performance.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_double, dp, dp]
    lib.grn_autotune.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ip]
    lib.grn_apply_tuning.argtypes = [vp, vp, ctypes.c_int]
    lib.grn_clear_tuning.argtypes = [vp]
    lib.grn_ceil_s.argtypes = [ctypes.c_double, ctypes.c_int]
    lib.grn_ceil_s.restype = ctypes.c_double
    lib.grn_model_parse.argtypes = [ctypes.c_char_p, vp, ctypes.c_int, ip, vp, ctypes.c_int, ip]
    lib.grn_model_format.argtypes = [ctypes.POINTER(_Model), ctypes.c_int, ctypes.c_char_p,
                                     ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    lib.grn_model_info.argtypes = [ctypes.POINTER(_Model), ctypes.c_int, ip, ip, dp]
    lib.grn_copy_probe.argtypes = [ctypes.c_long, dp]
    lib.grn_search_default_params.argtypes = [ctypes.POINTER(_SearchParams)]
    lib.grn_search_default_params.restype = None
    lib.grn_search_create.argtypes = [ctypes.POINTER(_SearchParams), vp, ctypes.POINTER(vp)]
    lib.grn_search_destroy.argtypes = [vp]
    lib.grn_search_destroy.restype = None
    lib.grn_search_pending.argtypes = [vp, ip, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                       ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.grn_search_advance.argtypes = [vp, vp, ctypes.POINTER(_GenerationStats)]
    lib.grn_search_run_generation.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(_GenerationStats)]
    lib.grn_search_best.argtypes = [vp, vp, ctypes.c_int, ip, vp, ctypes.c_int, ip, dp]
    lib.grn_search_pareto.argtypes = [vp, vp, ctypes.c_int, ip]
    lib.grn_search_pareto_model.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ip, vp, ctypes.c_int, ip]
    fp = ctypes.POINTER(ctypes.c_float)
    lib.grn_search_set_oversample.argtypes = [vp, ctypes.c_int]
    lib.grn_search_candidates.argtypes = [vp, ip, ctypes.POINTER(vp)]
    lib.grn_search_screen.argtypes = [vp, vp]
    lib.grn_search_set_screen_fraction.argtypes = [vp, ctypes.c_double]
    lib.grn_search_set_screen_policy.argtypes = [vp, ctypes.c_int]
    lib.grn_search_record_raw.argtypes = [vp, vp]
    lib.grn_search_run_generation_screened.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                       ctypes.POINTER(_GenerationStats),
                                                       ctypes.POINTER(_ScreenStats)]
    lib.grn_search_run_generation_coarse.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int,
                                                     ctypes.POINTER(_GenerationStats),
                                                     ctypes.POINTER(ctypes.c_double)]
    lib.grn_search_run_generation_scaled.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int,
                                                     ctypes.POINTER(_GenerationStats)]
    lib.grn_search_rescore.argtypes = [vp]
    lib.grn_surrogate_default_params.argtypes = [ctypes.POINTER(_SurrogateParams)]
    lib.grn_surrogate_default_params.restype = None
    lib.grn_surrogate_create.argtypes = [ctypes.c_int, ctypes.POINTER(_SurrogateParams), ctypes.POINTER(vp)]
    lib.grn_surrogate_destroy.argtypes = [vp]
    lib.grn_surrogate_destroy.restype = None
    lib.grn_surrogate_n_params.argtypes = []
    lib.grn_surrogate_n_params.restype = ctypes.c_long
    lib.grn_surrogate_encode.argtypes = [ctypes.POINTER(_Model), ctypes.c_int, vp]
    lib.grn_surrogate_add.argtypes = [vp, ctypes.c_int, vp, vp]
    lib.grn_surrogate_train.argtypes = [vp, ctypes.c_int, dp]
    lib.grn_surrogate_set_pair_weight.argtypes = [vp, ctypes.c_double]
    lib.grn_surrogate_predict.argtypes = [vp, ctypes.c_int, vp, vp]
    lib.grn_surrogate_predict_features.argtypes = [vp, ctypes.c_int, vp, vp]
    lib.grn_surrogate_gradient.argtypes = [vp, ctypes.c_int, vp, vp, vp, dp]
    lib.grn_surrogate_get_params.argtypes = [vp, vp]
    lib.grn_surrogate_set_params.argtypes = [vp, vp]
    lib.grn_surrogate_get_stats.argtypes = [vp, ctypes.POINTER(_SurrogateStats)]
    lib.grn_mlp_gemm_bf16.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp]
    for name in EXPORTED:
        if name not in ("grn_last_error", "grn_destroy", "grn_ceil_s", "grn_abi_version",
                        "grn_search_default_params", "grn_search_destroy", "grn_surrogate_default_params",
                        "grn_surrogate_destroy", "grn_surrogate_n_params"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise GrnError(rc, load_library().grn_last_error().decode(errors="replace"))


# ---- models -------------------------------------------------------------------

@dataclass
class Model:
    """Model (Src/Model/model.h): genes = ModelProd rows, links = ModelLink rows."""
    genes: np.ndarray = field(default_factory=lambda: np.zeros(0, GENE_DTYPE))
    links: np.ndarray = field(default_factory=lambda: np.zeros(0, LINK_DTYPE))

    def __post_init__(self):
        self.genes = np.ascontiguousarray(self.genes, dtype=GENE_DTYPE)
        self.links = np.ascontiguousarray(self.links, dtype=LINK_DTYPE)

    def _c(self) -> _Model:
        # the C side reads padded records: re-lay arrays that lost the layout
        # (np.concatenate of structured arrays can drop the padding)
        if self.genes.dtype != GENE_DTYPE or not self.genes.flags.c_contiguous:
            self.genes = np.ascontiguousarray(self.genes.astype(GENE_DTYPE))
        if self.links.dtype != LINK_DTYPE or not self.links.flags.c_contiguous:
            self.links = np.ascontiguousarray(self.links.astype(LINK_DTYPE))
        return _Model(len(self.genes), self.genes.ctypes.data if len(self.genes) else None,
                      len(self.links), self.links.ctypes.data if len(self.links) else None)

    @staticmethod
    def from_text(text: str) -> "Model":
        """Model::operator>> (Src/Model/model.cpp:659-706)."""
        lib = load_library()
        g = np.zeros(256, GENE_DTYPE)
        l = np.zeros(1024, LINK_DTYPE)
        ng, nl = ctypes.c_int(), ctypes.c_int()
        _check(lib.grn_model_parse(text.encode(), g.ctypes.data, len(g), ctypes.byref(ng),
                                   l.ctypes.data, len(l), ctypes.byref(nl)))
        return Model(g[:ng.value].copy(), l[:nl.value].copy())

    def to_text(self, precision: int = 6) -> str:
        """Model::operator<< (Src/Model/model.cpp:627-657)."""
        lib = load_library()
        cm = self._c()
        n = ctypes.c_long()
        lib.grn_model_format(ctypes.byref(cm), precision, None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(lib.grn_model_format(ctypes.byref(cm), precision, buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def info(self, n_morphogens: int):
        """(genes in use, Hill links, algorithmic FLOPs per cell-step)."""
        lib = load_library()
        cm = self._c()
        g, l, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib.grn_model_info(ctypes.byref(cm), n_morphogens, ctypes.byref(g), ctypes.byref(l),
                                  ctypes.byref(f)))
        return g.value, l.value, f.value


def decode_rgb(images: np.ndarray, n_planes: int) -> np.ndarray:
    """SimState::loadMorphologyImage (Src/Simulator/simstate.cpp:71-90) of
    images[n][H][W] (uint32 0xAARRGGBB, scan lines top to bottom) ->
    planes [n][n_planes][H][W] (float64): plane k is byte 2 - k / 255, image
    line j is plane row H - 1 - j."""
    img = np.ascontiguousarray(images, np.uint32)
    n, H, W = img.shape
    out = np.zeros((n, n_planes, H, W), np.float64)
    lib = load_library()
    for e in range(n):
        o = np.zeros((n_planes, H, W), np.float64)
        _check(lib.grn_decode_rgb(W, H, img[e].ctypes.data, n_planes, o.ctypes.data))
        out[e] = o
    return out


def ceil_s(n: float, digits: int) -> float:
    """MathAlgo::ceilS (Src/Common/mathalgo.h:55-62)."""
    return load_library().grn_ceil_s(n, digits)


@dataclass
class SimParams:
    """The SimParams fields of the path (Src/Simulator/simparams.h:25-43)."""
    dim_x: int
    dim_y: int
    n_input_morphogens: int
    n_target_morphogens: int
    num_sim_steps: int
    dt: float
    min_conc: float = 0.0
    min_conc_change: float = -1.0
    dist_error_threshold: float = 0.0
    error_precision: int = 0
    fitness_kind: int = FITNESS_MSE

    def _c(self) -> _SimParams:
        return _SimParams(self.dim_x, self.dim_y, self.n_input_morphogens, self.n_target_morphogens,
                          self.num_sim_steps, self.dt, self.min_conc, self.min_conc_change,
                          self.dist_error_threshold, self.error_precision, self.fitness_kind)


class StagedBatch:
    """A batch resident on the device (Context.stage_batch)."""

    def __init__(self, ctx, handle, n):
        self._ctx, self._h, self.n = ctx, handle, n
        self._lib = ctx._lib

    def run(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), RESULT_DTYPE)
        _check(self._lib.grn_run_staged(self._ctx._h, self._h, out.ctypes.data))
        return out[: self.n]

    def close(self):
        if self._h:
            self._lib.grn_staged_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ModelBatch:
    """Models marshalled once into the C array grn_evaluate_batch reads (the
    way a C++ caller keeps its grn_model array); the models must not change
    while the batch is in use."""

    def __init__(self, models):
        self.models = list(models)
        self.array = (_Model * max(len(self.models), 1))(*[m._c() for m in self.models])

    def __len__(self):
        return len(self.models)


class Context:
    """One device context: stream, uploaded experiment sets, device buffers."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.grn_create(device, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self.params: list[SimParams] = []
        self._staged = weakref.WeakSet()  # staged batches to free before the context

    def close(self):
        if self._h:
            for s in list(self._staged):
                s.close()
            self._lib.grn_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_dataset(self, params: SimParams, init: np.ndarray, target: np.ndarray) -> int:
        """init [n_exp][n_in][H][W], target [n_exp][n_tgt][H][W] (float64)."""
        init = np.ascontiguousarray(init, dtype=np.float64)
        target = np.ascontiguousarray(target, dtype=np.float64)
        n_exp = init.shape[0] if init.ndim == 4 else target.shape[0]
        cp = params._c()
        ds = ctypes.c_int()
        _check(self._lib.grn_add_dataset(self._h, ctypes.byref(cp), n_exp,
                                         init.ctypes.data if init.size else None,
                                         target.ctypes.data if target.size else None,
                                         ctypes.byref(ds)))
        self.params.append(params)
        return ds.value

    def add_dataset_rgb(self, params: SimParams, init_images: np.ndarray, target_images: np.ndarray) -> int:
        """Experiment images as the reference stores them (uint32 0xAARRGGBB,
        [n_exp][dim_y][dim_x], scan lines top to bottom): the input planes from
        init_images, the target planes from target_images (decode_rgb)."""
        ii = np.ascontiguousarray(init_images, dtype=np.uint32)
        ti = np.ascontiguousarray(target_images, dtype=np.uint32)
        n_exp = ii.shape[0] if ii.ndim == 3 else ti.shape[0]
        for a in (ii, ti):
            if a.size and a.shape != (n_exp, params.dim_y, params.dim_x):
                raise ValueError(f"images must be [n_exp][{params.dim_y}][{params.dim_x}]")
        cp = params._c()
        ds = ctypes.c_int()
        _check(self._lib.grn_add_dataset_rgb(self._h, ctypes.byref(cp), n_exp, ii.ctypes.data if ii.size else None,
                                             ti.ctypes.data if ti.size else None, ctypes.byref(ds)))
        self.params.append(params)
        return ds.value

    def evaluate(self, model: Model, max_error: float = float("inf"), dataset: int = 0) -> np.void:
        """Plain path: one model per launch, synchronous (the reference's pattern)."""
        cm = model._c()
        r = _Result()
        _check(self._lib.grn_evaluate(self._h, dataset, ctypes.byref(cm), max_error, ctypes.byref(r)))
        out = np.zeros(1, RESULT_DTYPE)
        ctypes.memmove(out.ctypes.data, ctypes.byref(r), RESULT_DTYPE.itemsize)
        return out[0]

    def evaluate_batch(self, models, datasets=None, max_errors=None, dt=None,
                       path: int = PATH_AUTO) -> np.ndarray:
        """Batched path: all models in persistent launches, one sync.
        `models` is a list of Model or a ModelBatch (marshalled once)."""
        n = len(models)
        cms = models.array if isinstance(models, ModelBatch) else (_Model * max(n, 1))(*[m._c() for m in models])
        ds = None if datasets is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(datasets, np.int32), (n,)))
        me = None if max_errors is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(max_errors, np.float64), (n,)))
        dts = None if dt is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(dt, np.float64), (n,)))
        out = np.zeros(max(n, 1), RESULT_DTYPE)
        _check(self._lib.grn_evaluate_batch(
            self._h, n, None if ds is None else ds.ctypes.data, ctypes.addressof(cms),
            None if me is None else me.ctypes.data, None if dts is None else dts.ctypes.data,
            path, out.ctypes.data))
        return out[:n]

    def evaluate_batch_noalloc(self, models, datasets=None, max_errors=None, dt=None) -> np.ndarray:
        """grn_evaluate_batch_noalloc: the on-chip batched path with every batch
        buffer in the device's fixed scratch array (at most NOALLOC_MAX_MODELS
        models; no device allocation)."""
        n = len(models)
        cms = models.array if isinstance(models, ModelBatch) else (_Model * max(n, 1))(*[m._c() for m in models])
        ds = None if datasets is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(datasets, np.int32), (n,)))
        me = None if max_errors is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(max_errors, np.float64), (n,)))
        dts = None if dt is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(dt, np.float64), (n,)))
        out = np.zeros(max(n, 1), RESULT_DTYPE)
        _check(self._lib.grn_evaluate_batch_noalloc(
            self._h, n, None if ds is None else ds.ctypes.data, ctypes.addressof(cms),
            None if me is None else me.ctypes.data, None if dts is None else dts.ctypes.data, out.ctypes.data))
        return out[:n]

    def stage_batch(self, models, datasets=None, max_errors=None, dt=None,
                    path: int = PATH_AUTO) -> "StagedBatch":
        """Compiles and uploads a batch once; StagedBatch.run() evaluates it
        with the programs resident on the device (same results as
        evaluate_batch on the same arguments)."""
        n = len(models)
        cms = models.array if isinstance(models, ModelBatch) else (_Model * max(n, 1))(*[m._c() for m in models])
        ds = None if datasets is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(datasets, np.int32), (n,)))
        me = None if max_errors is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(max_errors, np.float64), (n,)))
        dts = None if dt is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(dt, np.float64), (n,)))
        h = ctypes.c_void_p()
        _check(self._lib.grn_stage_batch(
            self._h, n, None if ds is None else ds.ctypes.data, ctypes.addressof(cms),
            None if me is None else me.ctypes.data, None if dts is None else dts.ctypes.data,
            path, ctypes.byref(h)))
        sb = StagedBatch(self, h.value, n)
        self._staged.add(sb)
        return sb

    def simulate_fields(self, model: Model, dataset: int = 0, experiment: int = 0,
                        dt: float = 0.0, path: int = PATH_AUTO) -> np.ndarray:
        """Final field [n_products][H][W] (float64) of one experiment."""
        p = self.params[dataset]
        cap = MAX_GENES * p.dim_x * p.dim_y
        buf = np.zeros(cap, np.float64)
        g = ctypes.c_int()
        cm = model._c()
        _check(self._lib.grn_simulate_fields(self._h, dataset, ctypes.byref(cm), dt, experiment, path,
                                             buf.ctypes.data, cap, ctypes.byref(g)))
        return buf[: g.value * p.dim_x * p.dim_y].reshape(g.value, p.dim_y, p.dim_x).copy()

    def test_performance(self, model_text: str, n: int, expected: float, dataset: int = 0):
        """ErrorCalculatorMultiThread::testPerformance: n copies in one batch,
        exact comparison.  Returns (ok, largest |fitness - expected|, kernel ms)."""
        dif, ms = ctypes.c_double(), ctypes.c_double()
        rc = self._lib.grn_test_performance(self._h, dataset, model_text.encode(), n, expected,
                                            ctypes.byref(dif), ctypes.byref(ms))
        if rc not in (0, -11):
            _check(rc)
        return rc == 0, dif.value, ms.value

    def set_launch_mode(self, mode: int):
        """LAUNCH_EAGER or LAUNCH_GRAPH (capture once per kernel-class layout, then replay)."""
        _check(self._lib.grn_set_launch_mode(self._h, mode))

    def autotune(self, models, datasets=None, sample_models: int = 512, step_cap: int = 200) -> list:
        """Times every kernel configuration of every size class in the batch
        (sample_models per class, step_cap steps) and keeps the fastest for
        later PATH_AUTO batches.  Returns one report dict per class."""
        n = len(models)
        cms = models.array if isinstance(models, ModelBatch) else (_Model * max(n, 1))(*[m._c() for m in models])
        ds = None if datasets is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(datasets, np.int32), (n,)))
        cap = 256
        rep = (_TuneClass * cap)()
        k = ctypes.c_int()
        _check(self._lib.grn_autotune(self._h, n, None if ds is None else ds.ctypes.data, ctypes.addressof(cms),
                                      sample_models, step_cap, ctypes.addressof(rep), cap, ctypes.byref(k)))
        return [_tune_to_dict(rep[i]) for i in range(min(k.value, cap))]

    def apply_tuning(self, reports: list):
        """Re-applies saved autotuner reports (each class's `chosen` candidate)."""
        arr = (_TuneClass * max(len(reports), 1))(*[_tune_from_dict(r) for r in reports])
        _check(self._lib.grn_apply_tuning(self._h, ctypes.addressof(arr), len(reports)))

    def clear_tuning(self):
        _check(self._lib.grn_clear_tuning(self._h))

    def stats(self) -> dict:
        s = _Stats()
        _check(self._lib.grn_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}


# ---- the search that calls the path ---------------------------------------------

def search_params(**overrides) -> dict:
    """The reference's search defaults (grn_search_default_params) with overrides."""
    p = _SearchParams()
    load_library().grn_search_default_params(ctypes.byref(p))
    out = {k: getattr(p, k) for k, _ in _SearchParams._fields_}
    for k, v in overrides.items():
        if k not in out:
            raise KeyError(k)
        out[k] = v
    return out


def _models_from_c(n: int, ptr: int) -> list:
    arr = (_Model * n).from_address(ptr) if n else []
    out = []
    for cm in arr:
        g = np.zeros(cm.n_genes, GENE_DTYPE)
        l = np.zeros(cm.n_links, LINK_DTYPE)
        if cm.n_genes:
            ctypes.memmove(g.ctypes.data, cm.genes, cm.n_genes * GENE_DTYPE.itemsize)
        if cm.n_links:
            ctypes.memmove(l.ctypes.data, cm.links, cm.n_links * LINK_DTYPE.itemsize)
        out.append(Model(g, l))
    return out


class Search:
    """Island search by deterministic crowding (Src/Search/search.cpp,
    searchalgodetcrowd.cpp): the caller of the simulation path."""

    def __init__(self, island_datasets=None, **params):
        lib = load_library()
        self._lib = lib
        self.params = search_params(**params)
        cp = _SearchParams(**self.params)
        ids = None
        if island_datasets is not None:
            ids = np.ascontiguousarray(np.broadcast_to(np.asarray(island_datasets, np.int32),
                                                       (self.params["n_islands"],)))
        h = ctypes.c_void_p()
        _check(lib.grn_search_create(ctypes.byref(cp), None if ids is None else ids.ctypes.data,
                                     ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            self._lib.grn_search_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_pending(self) -> int:
        n = ctypes.c_int()
        _check(self._lib.grn_search_pending(self._h, ctypes.byref(n), None, None, None, None))
        return n.value

    def pending(self):
        """(models, datasets, max_errors, dt) of the individuals awaiting an error."""
        n = ctypes.c_int()
        pm, pd, pe, pt = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(self._lib.grn_search_pending(self._h, ctypes.byref(n), ctypes.byref(pm), ctypes.byref(pd),
                                            ctypes.byref(pe), ctypes.byref(pt)))
        k = n.value
        models = _models_from_c(k, pm.value)

        def arr(ptr, ct, dtype):
            if not k:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array((ct * k).from_address(ptr.value)).astype(dtype)
        return (models, arr(pd, ctypes.c_int, np.int32), arr(pe, ctypes.c_double, np.float64),
                arr(pt, ctypes.c_double, np.float64))

    def advance(self, errors) -> dict:
        e = np.ascontiguousarray(errors, dtype=np.float64)
        st = _GenerationStats()
        _check(self._lib.grn_search_advance(self._h, e.ctypes.data if e.size else None, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _GenerationStats._fields_}

    def run_generation(self, ctx: "Context", path: int = PATH_AUTO) -> dict:
        st = _GenerationStats()
        _check(self._lib.grn_search_run_generation(self._h, ctx._h, path, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _GenerationStats._fields_}

    def set_oversample(self, k: int):
        """k candidates per offspring slot (1 = the reference's search)."""
        _check(self._lib.grn_search_set_oversample(self._h, k))

    def candidates(self) -> list:
        n = ctypes.c_int()
        pm = ctypes.c_void_p()
        _check(self._lib.grn_search_candidates(self._h, ctypes.byref(n), ctypes.byref(pm)))
        return _models_from_c(n.value, pm.value)

    def screen(self, scores):
        sc = np.ascontiguousarray(scores, dtype=np.float64)
        _check(self._lib.grn_search_screen(self._h, sc.ctypes.data))

    def set_screen_fraction(self, fraction: float):
        """Fraction of each island's slots screen() sends to the simulator:
        those whose kept candidate's score most undercuts the parent's
        recorded ln raw error (record_raw); the rest keep their parents."""
        _check(self._lib.grn_search_set_screen_fraction(self._h, float(fraction)))

    def set_screen_policy(self, policy: int):
        """SCREEN_ABSOLUTE (a candidate's ln error) or SCREEN_DELTA (its change
        of ln error from the slot parent) for run_generation_screened."""
        _check(self._lib.grn_search_set_screen_policy(self._h, int(policy)))

    def record_raw(self, raw_errors):
        """Raw errors of the pending models (before advance)."""
        r = np.ascontiguousarray(raw_errors, dtype=np.float64)
        if r.shape != (self.n_pending(),):
            raise ValueError(f"expected {self.n_pending()} raw errors")
        _check(self._lib.grn_search_record_raw(self._h, r.ctypes.data if r.size else None))

    def run_generation_screened(self, ctx: "Context", surrogate: "Surrogate", train_steps: int = 50,
                                path: int = PATH_AUTO):
        """(generation stats, screening stats) of one surrogate-screened generation."""
        st, sc = _GenerationStats(), _ScreenStats()
        _check(self._lib.grn_search_run_generation_screened(self._h, ctx._h, surrogate._h, train_steps, path,
                                                            ctypes.byref(st), ctypes.byref(sc)))
        return ({k: getattr(st, k) for k, _ in _GenerationStats._fields_},
                {k: getattr(sc, k) for k, _ in _ScreenStats._fields_})

    def run_generation_coarse(self, ctx: "Context", coarse_datasets, dif_scale: float, path: int = PATH_AUTO):
        """(generation stats, coarse-pass ms) of one generation whose candidates
        are screened by simulating them on coarse_datasets[island] (a coarser
        grid of the island's experiments) with their diffusion constants times
        dif_scale; see grn_search_run_generation_coarse."""
        ds = np.ascontiguousarray(coarse_datasets, dtype=np.int32)
        if ds.shape != (self.params["n_islands"],):
            raise ValueError(f"expected {self.params['n_islands']} coarse dataset ids")
        st, ms = _GenerationStats(), ctypes.c_double()
        _check(self._lib.grn_search_run_generation_coarse(self._h, ctx._h, ds.ctypes.data, float(dif_scale), path,
                                                          ctypes.byref(st), ctypes.byref(ms)))
        return {k: getattr(st, k) for k, _ in _GenerationStats._fields_}, ms.value

    def run_generation_scaled(self, ctx: "Context", island_datasets, dif_scale: float,
                              path: int = PATH_AUTO) -> dict:
        """One plain generation whose pending models are simulated on
        island_datasets[island] (say a coarser grid of the island's
        experiments) with their diffusion constants times dif_scale; see
        grn_search_run_generation_scaled."""
        ds = np.ascontiguousarray(island_datasets, dtype=np.int32)
        if ds.shape != (self.params["n_islands"],):
            raise ValueError(f"expected {self.params['n_islands']} dataset ids")
        st = _GenerationStats()
        _check(self._lib.grn_search_run_generation_scaled(self._h, ctx._h, ds.ctypes.data, float(dif_scale), path,
                                                          ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _GenerationStats._fields_}

    def rescore(self):
        """Every island member pending again, as an initial population (the
        next generation records their errors without selection): after a
        change of fidelity."""
        _check(self._lib.grn_search_rescore(self._h))

    def pareto(self) -> list:
        """The Pareto front on (error, complexity in use) as
        Search::recalculateParetoFront keeps it (search.cpp:348-377): one dict
        per member (error, complexity, generation, island, model)."""
        n = ctypes.c_int()
        _check(self._lib.grn_search_pareto(self._h, None, 0, ctypes.byref(n)))
        arr = (_ParetoEntry * max(n.value, 1))()
        _check(self._lib.grn_search_pareto(self._h, ctypes.addressof(arr), n.value, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            e = arr[i]
            g = np.zeros(max(e.n_genes, 1), GENE_DTYPE)
            l = np.zeros(max(e.n_links, 1), LINK_DTYPE)
            ng, nl = ctypes.c_int(), ctypes.c_int()
            _check(self._lib.grn_search_pareto_model(self._h, i, g.ctypes.data, len(g), ctypes.byref(ng),
                                                     l.ctypes.data, len(l), ctypes.byref(nl)))
            out.append({"error": e.error, "complexity": e.complexity, "generation": e.generation,
                        "island": e.island, "model": Model(g[:ng.value].copy(), l[:nl.value].copy())})
        return out

    def best(self):
        """(Model, error) of the lowest-error individual (ties: lower complexity)."""
        g = np.zeros(MAX_GENES * 16, GENE_DTYPE)
        l = np.zeros(MAX_LINKS * 16, LINK_DTYPE)
        ng, nl, e = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self._lib.grn_search_best(self._h, g.ctypes.data, len(g), ctypes.byref(ng), l.ctypes.data,
                                         len(l), ctypes.byref(nl), ctypes.byref(e)))
        return Model(g[:ng.value].copy(), l[:nl.value].copy()), e.value


# ---- the learned pre-screen -----------------------------------------------------

def surrogate_params(**overrides) -> dict:
    p = _SurrogateParams()
    load_library().grn_surrogate_default_params(ctypes.byref(p))
    out = {k: getattr(p, k) for k, _ in _SurrogateParams._fields_}
    for k, v in overrides.items():
        if k not in out:
            raise KeyError(k)
        out[k] = v
    return out


def encode(model: Model, n_morphogens: int) -> np.ndarray:
    """The surrogate's 468 features of a model."""
    f = np.zeros(MLP_FEATURES, np.float32)
    cm = model._c()
    _check(load_library().grn_surrogate_encode(ctypes.byref(cm), n_morphogens, f.ctypes.data))
    return f


def n_surrogate_params() -> int:
    return int(load_library().grn_surrogate_n_params())


def mlp_gemm_bf16(a_bits: np.ndarray, bt_bits: np.ndarray, device: int = 0) -> np.ndarray:
    """C = A . Bt^T on the surrogate's MFMA GEMM; operands as uint16 bf16 bit patterns."""
    a = np.ascontiguousarray(a_bits, np.uint16)
    bt = np.ascontiguousarray(bt_bits, np.uint16)
    M, K = a.shape
    N = bt.shape[0]
    c = np.zeros((M, N), np.float32)
    _check(load_library().grn_mlp_gemm_bf16(device, M, N, K, a.ctypes.data, bt.ctypes.data, c.ctypes.data))
    return c


def _model_array(models):
    n = len(models)
    return (_Model * max(n, 1))(*[m._c() for m in models])


class Surrogate:
    """MLP pre-screen (468 -> 4 x 512 GELU -> 1), BF16 MFMA kernels, Adam."""

    def __init__(self, device: int = 0, **params):
        lib = load_library()
        self._lib = lib
        self.params = surrogate_params(**params)
        cp = _SurrogateParams(**self.params)
        h = ctypes.c_void_p()
        _check(lib.grn_surrogate_create(device, ctypes.byref(cp), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            self._lib.grn_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, models, errors):
        e = np.ascontiguousarray(errors, np.float64)
        cms = _model_array(models)
        _check(self._lib.grn_surrogate_add(self._h, len(models), ctypes.addressof(cms), e.ctypes.data))

    def train(self, steps: int) -> float:
        loss = ctypes.c_double()
        _check(self._lib.grn_surrogate_train(self._h, steps, ctypes.byref(loss)))
        return loss.value

    def set_pair_weight(self, weight: float):
        """Pair mode (weight > 0, set while the store is empty): children are
        stored beside their slot parents and training also fits the predicted
        change of ln error from parent to child."""
        _check(self._lib.grn_surrogate_set_pair_weight(self._h, float(weight)))

    def predict(self, models) -> np.ndarray:
        out = np.zeros(max(len(models), 1), np.float64)
        cms = _model_array(models)
        _check(self._lib.grn_surrogate_predict(self._h, len(models), ctypes.addressof(cms), out.ctypes.data))
        return out[:len(models)]

    def predict_features(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        out = np.zeros(max(len(f), 1), np.float64)
        _check(self._lib.grn_surrogate_predict_features(self._h, len(f), f.ctypes.data, out.ctypes.data))
        return out[:len(f)]

    def gradient(self, features: np.ndarray, targets: np.ndarray):
        """(loss, flat gradient) of one forward/backward pass, no update."""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(targets, np.float64)
        g = np.zeros(n_surrogate_params(), np.float32)
        loss = ctypes.c_double()
        _check(self._lib.grn_surrogate_gradient(self._h, len(f), f.ctypes.data, t.ctypes.data, g.ctypes.data,
                                                ctypes.byref(loss)))
        return loss.value, g

    def get_params(self) -> np.ndarray:
        p = np.zeros(n_surrogate_params(), np.float32)
        _check(self._lib.grn_surrogate_get_params(self._h, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray):
        q = np.ascontiguousarray(p, np.float32)
        assert q.size == n_surrogate_params()
        _check(self._lib.grn_surrogate_set_params(self._h, q.ctypes.data))

    def stats(self) -> dict:
        s = _SurrogateStats()
        _check(self._lib.grn_surrogate_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _SurrogateStats._fields_}


def copy_probe(nbytes: int) -> float:
    """Runs the calibration copy kernel over nbytes; returns its time in ms."""
    ms = ctypes.c_double()
    _check(load_library().grn_copy_probe(nbytes, ctypes.byref(ms)))
    return ms.value


def device_count() -> int:
    n = ctypes.c_int()
    _check(load_library().grn_device_count(ctypes.byref(n)))
    return n.value
