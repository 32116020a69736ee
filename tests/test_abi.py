"""CPU tests of the C ABI library: it loads, exports every entry point that
include/grn.h declares, and its host logic (model text format, model
compiler bookkeeping, ceilS) behaves.  No compute call is made here."""
import os
import re
import subprocess

import numpy as np
import pytest

import grn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "grn.h")
KAT = open(os.path.join(ROOT, "tests", "golden", "kat_model.txt")).read().strip()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(grn_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(grn.LIB_PATH), "build libgrn_hip.so first"
    out = subprocess.run(["nm", "-D", "--defined-only", grn.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (grn_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert set(grn.EXPORTED) == set(declared_functions())


def test_library_loads_and_reports_version():
    lib = grn.load_library()
    assert lib.grn_abi_version() == 3
    assert lib.grn_last_error() is not None


def test_ceil_s_matches_reference_values():
    assert grn.ceil_s(0.0, 4) == 0.0
    assert grn.ceil_s(1.23456, 3) == pytest.approx(1.24, abs=0)
    assert grn.ceil_s(-1.23456, 3) == pytest.approx(-1.23, abs=0)
    once = grn.ceil_s(0.048701011019999997, 10)
    assert once == pytest.approx(0.04870101102, rel=1e-15)
    assert grn.ceil_s(once, 10) == once


def test_kat_model_parses_and_round_trips():
    m = grn.Model.from_text(KAT)
    assert len(m.genes) == 13 and len(m.links) == 19
    assert set(m.genes["label"]) >= {0, 1}
    assert m.genes["init"][m.genes["label"] == 0][0] == -1  # red morphogen tag
    assert m.genes["init"][m.genes["label"] == 1][0] == -2  # green morphogen tag
    text = m.to_text()
    assert text == KAT  # six significant digits, as the reference writes them
    again = grn.Model.from_text(text)
    assert np.array_equal(again.genes, m.genes) and np.array_equal(again.links, m.links)
    full = grn.Model.from_text(m.to_text(precision=17))
    assert np.array_equal(full.genes, m.genes)


def test_kat_model_in_use_and_flop_count():
    m = grn.Model.from_text(KAT)
    G, L, F = m.info(n_morphogens=2)
    # in use: 0, 1, 64653, 107652 (the reference calls it a 4-product model)
    assert G == 4
    # Hill links into in-use genes with an activator: gene 0 <- 64653 (act);
    # gene 1 <- 0 (act), 1 (act), 107652 (rep); 64653 <- 0 (act);
    # 107652 <- 1 (act), 64653 (rep)
    assert L == 7
    G_h, D = 4, 1  # every in-use gene has an activator; only label 0 diffuses
    assert F == 5 * L + (L - G_h) + 8 * G + 7 * D


@pytest.mark.parametrize("bad", ["", "(", "(0 1 2 3|)", "(0 0 1 1 0*0 0 1)", "0 0 1 1 0*)"])
def test_malformed_text_is_rejected(bad):
    with pytest.raises(grn.GrnError):
        grn.Model.from_text(bad)


def test_model_without_links_parses():
    m = grn.Model.from_text("(0 -1 0.5 0.2 0|1 -2 0.1 0.3 0.05*)")
    assert len(m.genes) == 2 and len(m.links) == 0
    assert m.to_text() == "(0 -1 0.5 0.2 0|1 -2 0.1 0.3 0.05*)"


def test_compile_rejects_missing_product_and_capacity():
    # label 7 regulates gene 0 but has no product
    m = grn.Model.from_text("(0 0 0.5 0.2 0|1 0 0.1 0.3 0*7 0 0.5 2 0)")
    with pytest.raises(grn.GrnError) as e:
        m.info(2)
    assert e.value.code == -2
    # 17 genes in use exceed the kernels' 16
    genes = np.zeros(17, grn.GENE_DTYPE)
    genes["label"] = np.arange(17)
    links = np.zeros(16, grn.LINK_DTYPE)
    links["regulator"] = np.arange(1, 17)
    links["regulated"] = 0
    links["dis_const"] = 0.5
    links["hill_coef"] = 1.0
    with pytest.raises(grn.GrnError) as e:
        grn.Model(genes, links).info(1)
    assert e.value.code == -3


def _c_layout(tmp_path, structs):
    """sizeof and field offsets of the header's structs, as gcc lays them out."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "grn.h"', "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_struct_layouts_match_header(tmp_path):
    import ctypes
    mirrors = {"grn_sim_params": grn._SimParams, "grn_model": grn._Model, "grn_result": grn._Result,
               "grn_stats": grn._Stats, "grn_search_params": grn._SearchParams,
               "grn_generation_stats": grn._GenerationStats, "grn_surrogate_params": grn._SurrogateParams,
               "grn_surrogate_stats": grn._SurrogateStats, "grn_screen_stats": grn._ScreenStats,
               "grn_tune_candidate": grn._TuneCandidate, "grn_tune_class": grn._TuneClass,
               "grn_pareto_entry": grn._ParetoEntry}
    dtypes = {"grn_gene": grn.GENE_DTYPE, "grn_link": grn.LINK_DTYPE}
    structs = {k: [f for f, _ in v._fields_] for k, v in mirrors.items()}
    structs.update({k: list(v.names) for k, v in dtypes.items()})
    lay = _c_layout(tmp_path, structs)
    for name, cls in mirrors.items():
        assert lay[name] == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert lay[f"{name}.{f}"] == getattr(cls, f).offset, f"{name}.{f}"
    for name, dt in dtypes.items():
        assert lay[name] == dt.itemsize, name
        for f in dt.names:
            assert lay[f"{name}.{f}"] == dt.fields[f][1], f"{name}.{f}"
    assert grn.RESULT_DTYPE.itemsize == lay["grn_result"]
    for f in grn.RESULT_DTYPE.names:
        assert grn.RESULT_DTYPE.fields[f][1] == lay[f"grn_result.{f}"]


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message, before
    touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_coarse_schedule_arguments():
    """--coarse-generations: a count on --coarse-side or side:count phases;
    evolve() refuses negative phases and schedules that leave no full-size
    generation (checked before any device work)."""
    import argparse
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    import evolve
    assert bench.coarse_schedule(argparse.Namespace(coarse_generations="100")) == 100
    assert bench.coarse_schedule(argparse.Namespace(coarse_generations="8:100,16:50")) == [(8, 100), (16, 50)]
    prob = evolve.scaled_config4(seed=0, n_islands=1, island_size=8)
    for bad in (-1, 5, [(8, 3), (16, 2)], [(8, -1)]):
        with pytest.raises(ValueError):
            evolve.evolve(None, prob, False, max_generations=5, threshold=0.0, coarse_generations=bad)


def test_decode_rgb_matches_oracle_and_known_pixels():
    """grn_decode_rgb (the library's SimState::loadMorphologyImage) against
    the oracle's restatement and against pixels with known channel bytes."""
    import grn_oracle as ora
    rng = np.random.default_rng(4)
    img = rng.integers(0, 2 ** 32, (2, 5, 7), dtype=np.uint64).astype(np.uint32)
    for n_planes in (1, 2, 3):
        np.testing.assert_array_equal(grn.decode_rgb(img, n_planes), ora.decode_rgb(img, n_planes))
    px = np.array([[[0xFF0A1428, 0x00000000], [0x00FFFFFF, 0x12345678]]], np.uint32)
    out = grn.decode_rgb(px, 3)
    np.testing.assert_array_equal(out[0, :, 1, 0], [10 / 255, 20 / 255, 40 / 255])  # top line -> last row
    np.testing.assert_array_equal(out[0, :, 0, 0], [1.0, 1.0, 1.0])
    np.testing.assert_array_equal(out[0, :, 0, 1], [0x34 / 255, 0x56 / 255, 0x78 / 255])
    with pytest.raises(grn.GrnError):
        grn.decode_rgb(px, 4)
