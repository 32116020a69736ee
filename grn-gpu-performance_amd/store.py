"""Results store: a search's parameters, per-generation statistics, Pareto
front and best models in an SQLite file laid out as the reference persists
them (Src/DB/db.cpp opens an SQLite database through QSQLITE; every element
submits its row with DBElementData::submit).

Table and column names follow the reference's submit() calls:
  SimParams        Src/Simulator/simparams.cpp:77-101
  SearchParams     Src/Search/searchparams.cpp:53-61 (NumGenerations as load() reads it, :45)
  Search           Src/Search/search.cpp:436-451 (SearchParams and SimParams are references)
  Deme             Src/Search/deme.cpp:58-65 (a row per deme, referencing its Search)
  Generation       Src/Search/generation.cpp:96-112 (referencing its Deme)
  Individual       Src/Search/individual.cpp:117-126; its Model column holds the
                   model in the reference's text format (Model::operator<<,
                   model.cpp:627-657), so a stored model loads in either program
  GenerationIndividual  Src/Search/generationindividual.cpp:49-60
Every table has the integer primary key Id that DBElementData joins on.

The islands advance in lockstep here (one batched evaluation per
generation), so a search stores one Deme row per island and its generation
statistics -- Generation::calcPopulSta over all islands -- under the first
island's Deme; the Pareto front (Search::submitParetoFront, search.cpp:461-466)
is stored as Individual rows linked to the last stored generation, with the
front position as GenerationIndividual.Rank.
"""
from __future__ import annotations

import datetime
import sqlite3

import grn

DATETIME_FORMAT = "%Y-%m-%dT%H:%M:%S.%fZ"  # DB::DatetimeFormat (db.cpp:18), to microseconds

SCHEMA = {
    "SimParams": ["Name TEXT", "DimX INTEGER", "DimY INTEGER", "Dt REAL", "Dx REAL", "NumSimSteps INTEGER",
                  "NumSims INTEGER", "NumInputMorphogens INTEGER", "NumTargetMorphogens INTEGER",
                  "NumMaxProducts INTEGER", "NumMaxLinks INTEGER", "MinConc REAL", "MinConcChange REAL",
                  "DistErrorThreshold REAL", "ErrorPrecision INTEGER", "NumCPUSlaves INTEGER",
                  "NumGPUSlaves INTEGER", "NumGPUThreads INTEGER", "NumTestIndividuals INTEGER",
                  "MultiGPU INTEGER"],
    "SearchParams": ["Name TEXT", "NumDemes INTEGER", "DemesSize INTEGER", "NumGenerations INTEGER",
                     "MigrationPeriod INTEGER", "SaveIndividuals INTEGER", "TestModel INTEGER"],
    "Search": ["Name TEXT", "RandSeed INTEGER", "StartDatetime TEXT", "EndDatetime TEXT",
               "SearchParams INTEGER REFERENCES SearchParams(Id)", "SimParams INTEGER REFERENCES SimParams(Id)"],
    "Deme": ["Search INTEGER REFERENCES Search(Id)"],
    "Generation": ["Deme INTEGER REFERENCES Deme(Id)", "Ind INTEGER", "Time REAL", "MinError REAL",
                   "MeanError REAL", "MaxError REAL", "MinComp REAL", "MeanComp REAL", "MaxComp REAL",
                   "BestComp INTEGER"],
    "Individual": ["Search INTEGER REFERENCES Search(Id)", "Model TEXT", "Complexity INTEGER", "Error REAL",
                   "SimTime REAL", "Parent1 INTEGER", "Parent2 INTEGER"],
    "GenerationIndividual": ["Generation INTEGER REFERENCES Generation(Id)",
                             "Individual INTEGER REFERENCES Individual(Id)", "Rank INTEGER", "CrowdDist REAL"],
}


def _now() -> str:
    return datetime.datetime.now(datetime.timezone.utc).strftime(DATETIME_FORMAT)


class Store:
    """An SQLite results file (created with the tables above if missing)."""

    def __init__(self, path: str):
        self.db = sqlite3.connect(path)
        self.db.execute("PRAGMA foreign_keys = true;")  # DB::connect's normal mode (db.cpp:52-55)
        for table, cols in SCHEMA.items():
            self.db.execute(f"CREATE TABLE IF NOT EXISTS {table} (Id INTEGER PRIMARY KEY AUTOINCREMENT, "
                            + ", ".join(cols) + ")")
        self.db.commit()

    def close(self):
        self.db.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _insert(self, table: str, values: dict) -> int:
        cols = list(values)
        cur = self.db.execute(f"INSERT INTO {table} ({', '.join(cols)}) VALUES ({', '.join('?' * len(cols))})",
                              [values[c] for c in cols])
        return int(cur.lastrowid)

    # ---- parameters ------------------------------------------------------------
    def add_sim_params(self, p: grn.SimParams, name: str = "", max_products: int = -1, max_links: int = -1,
                       dx: float = 1.0) -> int:
        """SimParams::submit (simparams.cpp:77-101)."""
        return self._insert("SimParams", {
            "Name": name, "DimX": p.dim_x, "DimY": p.dim_y, "Dt": p.dt, "Dx": dx, "NumSimSteps": p.num_sim_steps,
            "NumSims": 1, "NumInputMorphogens": p.n_input_morphogens,
            "NumTargetMorphogens": p.n_target_morphogens, "NumMaxProducts": max_products,
            "NumMaxLinks": max_links, "MinConc": p.min_conc, "MinConcChange": p.min_conc_change,
            "DistErrorThreshold": p.dist_error_threshold, "ErrorPrecision": p.error_precision,
            "NumCPUSlaves": 0, "NumGPUSlaves": 1, "NumGPUThreads": 0, "NumTestIndividuals": 0, "MultiGPU": 0})

    def add_search_params(self, search: dict, name: str = "", save_individuals: int = 2) -> int:
        """SearchParams::submit (searchparams.cpp:53-61), from grn.search_params()."""
        return self._insert("SearchParams", {
            "Name": name, "NumDemes": search["n_islands"], "DemesSize": search["island_size"],
            "NumGenerations": search.get("n_generations", 0), "MigrationPeriod": search["migration_period"],
            "SaveIndividuals": save_individuals, "TestModel": 0})

    # ---- a search --------------------------------------------------------------
    def begin_search(self, name: str, seed: int, sim_params_id: int, search_params_id: int,
                     n_islands: int) -> int:
        """Search::submit (search.cpp:436-451) at the start of a run, with one
        Deme row per island."""
        sid = self._insert("Search", {"Name": name, "RandSeed": int(seed), "StartDatetime": _now(),
                                      "EndDatetime": None, "SearchParams": search_params_id,
                                      "SimParams": sim_params_id})
        for _ in range(n_islands):
            self._insert("Deme", {"Search": sid})
        self.db.commit()
        return sid

    def end_search(self, search_id: int):
        self.db.execute("UPDATE Search SET EndDatetime = ? WHERE Id = ?", (_now(), search_id))
        self.db.commit()

    def _first_deme(self, search_id: int) -> int:
        return int(self.db.execute("SELECT MIN(Id) FROM Deme WHERE Search = ?", (search_id,)).fetchone()[0])

    def add_generation(self, search_id: int, stats: dict, time_s: float) -> int:
        """Generation::submit (generation.cpp:96-112) with the statistics
        grn_search_advance returns (Generation::calcPopulSta over all islands;
        the complexity columns hold the best individual's complexity)."""
        gid = self._insert("Generation", {
            "Deme": self._first_deme(search_id), "Ind": int(stats["generation"]), "Time": float(time_s),
            "MinError": stats["min_fit"], "MeanError": stats["mean_fit"], "MaxError": stats["max_fit"],
            "MinComp": stats["best_complexity"], "MeanComp": None, "MaxComp": None,
            "BestComp": stats["best_complexity"]})
        self.db.commit()
        return gid

    def add_individual(self, search_id: int, model: grn.Model, complexity: int, error: float,
                       sim_time: float = 0.0) -> int:
        """Individual::submit (individual.cpp:117-126): the model as text."""
        return self._insert("Individual", {"Search": search_id, "Model": model.to_text(17),
                                           "Complexity": int(complexity), "Error": float(error),
                                           "SimTime": float(sim_time), "Parent1": None, "Parent2": None})

    def add_pareto_front(self, search_id: int, front: list) -> list:
        """Search::submitParetoFront (search.cpp:461-466): every member of
        grn.Search.pareto() as an Individual, linked to the last stored
        generation with its front position as Rank."""
        row = self.db.execute("SELECT MAX(Generation.Id) FROM Generation JOIN Deme ON Generation.Deme = Deme.Id "
                              "WHERE Deme.Search = ?", (search_id,)).fetchone()
        gen = row[0]
        ids = []
        for rank, x in enumerate(front):
            iid = self.add_individual(search_id, x["model"], x["complexity"], x["error"])
            if gen is not None:
                self._insert("GenerationIndividual", {"Generation": gen, "Individual": iid, "Rank": rank,
                                                      "CrowdDist": None})
            ids.append(iid)
        self.db.commit()
        return ids

    # ---- reading back --------------------------------------------------------------
    def generations(self, search_id: int) -> list:
        rows = self.db.execute("SELECT Ind, Time, MinError, MeanError, MaxError, BestComp FROM Generation "
                               "JOIN Deme ON Generation.Deme = Deme.Id WHERE Deme.Search = ? ORDER BY Ind",
                               (search_id,)).fetchall()
        return [dict(zip(("generation", "time", "min_fit", "mean_fit", "max_fit", "best_complexity"), r))
                for r in rows]

    def pareto_front(self, search_id: int) -> list:
        """The stored front, in rank order: (Model, error, complexity)."""
        rows = self.db.execute(
            "SELECT Individual.Model, Individual.Error, Individual.Complexity FROM GenerationIndividual "
            "JOIN Individual ON GenerationIndividual.Individual = Individual.Id WHERE Individual.Search = ? "
            "AND GenerationIndividual.Generation = (SELECT MAX(Generation) FROM GenerationIndividual "
            "JOIN Individual ON GenerationIndividual.Individual = Individual.Id WHERE Individual.Search = ?) "
            "ORDER BY GenerationIndividual.Rank", (search_id, search_id)).fetchall()
        return [(grn.Model.from_text(m), e, c) for m, e, c in rows]

    def individual(self, individual_id: int):
        m, e, c = self.db.execute("SELECT Model, Error, Complexity FROM Individual WHERE Id = ?",
                                  (individual_id,)).fetchone()
        return grn.Model.from_text(m), e, c
