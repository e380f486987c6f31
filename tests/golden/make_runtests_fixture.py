"""Write runtests_3x2.json: the only known-answer data the reference's tests
hold for this path, transcribed from test/runtests.jl:5-16 (input matrix
``d`` with row 2 set to one, ``dshift1_expected`` and ``fdr_expected`` for
rates -1:1).  Matrices are stored row by row as in the Julia literal.

Run: python make_runtests_fixture.py  (rewrites the JSON beside it)
"""
import json
from pathlib import Path

FIXTURE = {
    "source": "test/runtests.jl:5-16,19-31",
    "Nf": 3, "Nt": 2,
    "d": [[0, 0], [1, 1], [0, 0]],
    "intshift_rate": 1,
    "dshift1_expected": [[0, 1], [1, 0], [0, 0]],
    "rates": {"first": -1, "step": 1, "length": 3},
    "fdr_expected": [[0, 0, 1], [1, 2, 1], [1, 0, 0]],
    "fftfdr_scale_in_test": 3,
    "zdtfdr_scale_in_test": 12,
    "Nl": 4,
}

if __name__ == "__main__":
    out = Path(__file__).with_name("runtests_3x2.json")
    out.write_text(json.dumps(FIXTURE, indent=1) + "\n")
    print(out)
