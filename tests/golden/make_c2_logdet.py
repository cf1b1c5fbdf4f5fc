"""Oracle log-determinant of the benchmark matrix (N=90,000 points, NB=900,
maxrank 150, eps 1e-8, statistics-2d-sqexp, beta 0.1, nugget 1e-2, seed 0),
written to c2_logdet.json.  About 35 s on 8 cores and 12 GB of memory."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import tlr_oracle as O  # noqa: E402

o = O.Oracle(90000, 900, 2, 1.0, 0.1, 1e-2, 0)
assert o.generate(1e-8, 150, 1) == 0
band = O.band_autotune(o.rank_gather(), o.NT, o.NB)
if band != 1:
    o.regenerate(band)
assert o.potrf(1e-8, 0, 150) == 0
json.dump({"N": 90000, "NB": 900, "maxrank": 150, "eps": 1e-8, "band": band, "logdet": o.logdet()},
          open(os.path.join(HERE, "c2_logdet.json"), "w"), indent=1)
