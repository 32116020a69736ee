"""The results store (grn-gpu-performance_amd/store.py): tables and columns as
the reference's elements submit them (Src/DB, Src/Search/*::submit), and a
search's generations and Pareto front read back exactly -- models through the
reference's text format.  CPU only: the search's errors come from a
deterministic fake fitness."""
import sqlite3

import numpy as np

import grn
import store
from test_search import coarse_error, complexity_in_use, drive


def test_store_round_trip(tmp_path):
    path = str(tmp_path / "results.db")
    p = grn.SimParams(32, 32, 2, 1, 1000, 0.01, fitness_kind=grn.FITNESS_REF, min_conc_change=1e-4,
                      error_precision=6)
    s = grn.Search(n_islands=3, island_size=16, n_morphogens=3, seed=4, n_generations=50)
    with store.Store(path) as st:
        sim_id = st.add_sim_params(p, name="gradients")
        sp_id = st.add_search_params(s.params, name="detcrowd")
        sid = st.begin_search("run 1", s.params["seed"], sim_id, sp_id, s.params["n_islands"])
        hist = drive(s, 8, fitness=coarse_error)
        for k, (stats, _, _) in enumerate(hist):
            st.add_generation(sid, stats, time_s=0.1 * k)
        front = s.pareto()
        st.add_pareto_front(sid, front)
        st.end_search(sid)
    # reopened: the reference's tables and columns are there
    db = sqlite3.connect(path)
    cols = {t: [r[1] for r in db.execute(f"pragma table_info({t})")] for t in store.SCHEMA}
    assert {"Model", "Complexity", "Error", "SimTime", "Parent1", "Parent2"} <= set(cols["Individual"])
    assert {"Ind", "MinError", "MeanError", "MaxError", "BestComp"} <= set(cols["Generation"])
    assert {"NumInputMorphogens", "MinConcChange", "ErrorPrecision"} <= set(cols["SimParams"])
    assert db.execute("SELECT COUNT(*) FROM Deme").fetchone()[0] == 3
    assert db.execute("SELECT EndDatetime FROM Search").fetchone()[0] is not None
    with store.Store(path) as st:
        gens = st.generations(sid)
        assert [g["generation"] for g in gens] == [h[0]["generation"] for h in hist]
        assert [g["min_fit"] for g in gens] == [h[0]["min_fit"] for h in hist]
        back = st.pareto_front(sid)
        assert len(back) == len(front)
        for (m, e, c), x in zip(back, front):
            assert m.to_text(17) == x["model"].to_text(17)
            assert e == x["error"] and c == x["complexity"]
            assert coarse_error(m) == e and complexity_in_use(m, 3) == c
