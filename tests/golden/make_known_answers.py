"""Writes known_answers.json: the known-answer cases that the reference
itself records for the helpers on the step's path (inputs and the expected
outputs stated in its comments).  Data only.

  sample6.f90:38-48   tdma on a 4x4 tridiagonal system -> 1, 2, 3, 4
  sample2.f90:142-145 gauss 3x3 -> 6, 17, 11
  sample2.f90:152-156 gauss 4x4 -> 1, -1, 2, -2
  sample2.f90:158-164 gauss 6x6 -> 0.5, 1, 1.5, 2, 2.5, -2.5
  sample5.f90:9-20    dsyev of [[2,1,1],[1,2,1],[1,1,2]] -> 1, 1, 4
"""
import json
import os

CASES = {
    "tdma_4": {"source": "sample6.f90:38-48",
               "a": [2.0, 2.0, 2.0, 2.0], "b": [1.0, 1.0, 1.0, 0.0], "c": [0.0, 1.0, 1.0, 1.0],
               "d": [4.0, 8.0, 12.0, 11.0], "x": [1.0, 2.0, 3.0, 4.0]},
    "gauss_3": {"source": "sample2.f90:142-145",
                "A": [[1.0, -2.0, 3.0], [2.0, -1.0, 1.0], [1.0, 3.0, -5.0]], "b": [5.0, 6.0, 2.0],
                "x": [6.0, 17.0, 11.0]},
    "gauss_4": {"source": "sample2.f90:152-156",
                "A": [[1.0, 1.0, 1.0, 1.0], [1.0, 1.0, 1.0, -1.0], [1.0, 1.0, -1.0, 1.0], [1.0, -1.0, 1.0, 1.0]],
                "b": [0.0, 4.0, -4.0, 2.0], "x": [1.0, -1.0, 2.0, -2.0]},
    "gauss_6": {"source": "sample2.f90:158-164",
                "A": [[1.0, 1.0, 1.0, 1.0, -1.0, 1.0], [1.0, 1.0, 1.0, -1.0, 1.0, 1.0],
                      [1.0, 1.0, -1.0, 1.0, 1.0, 1.0], [1.0, -1.0, 1.0, 1.0, 1.0, 1.0],
                      [-1.0, 1.0, 1.0, 1.0, 1.0, 1.0], [1.0, 1.0, 1.0, 1.0, 1.0, 1.0]],
                "b": [0.0, 1.0, 2.0, 3.0, 4.0, 5.0], "x": [0.5, 1.0, 1.5, 2.0, 2.5, -2.5]},
    "dsyev_3": {"source": "sample5.f90:9-20",
                "C6": [2.0, 1.0, 1.0, 2.0, 1.0, 2.0], "eig": [1.0, 1.0, 4.0]},
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(out, "w") as f:
        json.dump(CASES, f, indent=1)
    print(out)
