"""Write tdma_kat.json: the two known-answer tridiagonal systems of the
reference's own tests, Test_TDMA_noncyclic (src/basics_algorithms.f90:111-156,
data :123-127) and Test_TDMA_cyclic (:158-188, reference solution :179-181).
The numbers are data transcribed from those tests: the coefficient rows a, b,
c, the right-hand side r and the exact solutions (integers for the
non-cyclic system; rationals p/q for the cyclic one, stored as [p, q]).

    python tests/golden/make_tdma_kat.py
"""
import json
import os

A = [3.0, 1.0, 1.0, 7.0, 6.0, 3.0, 8.0, 6.0, 5.0, 4.0]
B = [2.0, 3.0, 3.0, 2.0, 2.0, 4.0, 1.0, 2.0, 4.0, 5.0]
C = [1.0, 2.0, 1.0, 6.0, 1.0, 3.0, 5.0, 7.0, 3.0, 5.0]
R = [1.0, 2.0, 6.0, 34.0, 10.0, 1.0, 4.0, 22.0, 25.0, 3.0]
NONCYCLIC = [1, -1, 2, 1, 3, -2, 0, 4, 2, -1]
CYCLIC = [[518663, 174746], [-299297, 174746], [182180, 87373], [5419, 3718], [480243, 174746],
          [-370592, 87373], [566251, 174746], [1212441, 174746], [-76, 47], [-187761, 174746]]

if __name__ == "__main__":
    out = {
        "source": "src/basics_algorithms.f90:111-188",
        "a": A, "b": B, "c": C, "r": R,
        "noncyclic_solution": NONCYCLIC,
        "cyclic_solution_rational": CYCLIC,
        "cyclic_note": "Preprocess_TDMA_coeffs on the first n-1 rows, Solve_TDMA(.true.) on n",
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tdma_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
