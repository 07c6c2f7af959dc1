"""Write ``tests/golden/fixtures.json`` from the reference's own test data.

Run once in a container where the reference is mounted at
``/root/reference`` (it is read as text; nothing there is imported or run):

    python tests/golden/make_fixtures.py

Harvested (data only):
  * literal Hamiltonians of ``tests/test_operators.py:39-88``
    (RandomHeisenberg n=2,3 h=0; XXZ n=2,3 delta=0.5)
  * ED eigenvalues / eigenvectors of ``tests/test_exact_diagonalization.py:15-42``
  * the 64 printed ED eigenvalues of RandomHeisenberg(n=6, h=3.0, seed=2021)
    in ``docs/source/notebooks/tsdrg.ipynb:140-150`` (8 significant digits)
Computed here independently of the oracle:
  * closed-form ground energies of the open XX chain (free fermions) and of
    the ferromagnetic XXZ chain at delta=1 (tnpy's sign convention)
  * brute-force sparse-Lanczos ground energies for XXZ(n=16, delta=0.5)
    and XXZ(n=10, delta=0.5) from bond-term Kronecker products
  * seeded RandomState field vectors
"""
from __future__ import annotations

import ast
import json
import math
import re
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "fixtures.json"


def _array_literals(text: str):
    """Every ``np.array([...])`` literal in a source text, as nested lists
    (``np.sqrt(2)`` is evaluated)."""
    out = []
    for m in re.finditer(r"np\.array\(\s*(\[)", text):
        start = m.start(1)
        depth = 0
        for i in range(start, len(text)):
            if text[i] == "[":
                depth += 1
            elif text[i] == "]":
                depth -= 1
                if depth == 0:
                    lit = text[start : i + 1]
                    break
        lit = lit.replace("np.sqrt(2)", repr(math.sqrt(2)))
        lit = re.sub(r"#.*", "", lit)
        try:
            out.append(_eval_numeric(lit))
        except (ValueError, SyntaxError):
            continue  # not a numeric literal (e.g. a comprehension)
    return out


def _eval_numeric(lit: str):
    # literal lists with simple arithmetic (1 / sqrt2, -1 / sqrt2)
    node = ast.parse(lit, mode="eval")
    allowed = (ast.Expression, ast.List, ast.Constant, ast.UnaryOp, ast.USub, ast.BinOp, ast.Div, ast.Load)
    for n in ast.walk(node):
        if not isinstance(n, allowed):
            raise ValueError(f"unexpected syntax in literal: {type(n).__name__}")
    return eval(compile(node, "<literal>", "eval"), {"__builtins__": {}})


def harvest_operator_literals():
    text = (REF / "tests/test_operators.py").read_text()
    arrays = _array_literals(text)
    # order in the file: two identity checks of SpinOperators, then the four
    # Hamiltonians of TestFullHamiltonian
    mats = [a for a in arrays if isinstance(a, list) and len(a) in (4, 8) and isinstance(a[0], list) and len(a[0]) == len(a)]
    assert len(mats) == 4, f"expected 4 Hamiltonian literals, found {len(mats)}"
    return {
        "random_heisenberg_n2_h0": mats[0],
        "random_heisenberg_n3_h0": mats[1],
        "xxz_n2_delta0.5": mats[2],
        "xxz_n3_delta0.5": mats[3],
    }


def harvest_ed_literals():
    text = (REF / "tests/test_exact_diagonalization.py").read_text()
    arrays = _array_literals(text)
    return {
        "heisenberg_n2_matrix": arrays[0],
        "heisenberg_n2_evals": arrays[1],
        "heisenberg_n2_evecs": arrays[2],
    }


def harvest_notebook_evals():
    nb = json.loads((REF / "docs/source/notebooks/tsdrg.ipynb").read_text())
    for cell in nb["cells"]:
        src = "".join(cell.get("source", []))
        if "RandomHeisenberg(N=6, h=3.0, seed=2021)" in src and "print(ed.evals)" in src:
            text = "".join(cell["outputs"][0]["text"])
            vals = [float(x) for x in text.replace("[", " ").replace("]", " ").split()]
            assert len(vals) == 64
            return vals
    raise RuntimeError("notebook cell not found")


def xx_closed_form(n: int) -> float:
    eps = -np.cos(np.arange(1, n + 1) * math.pi / (n + 1))
    return float(eps[eps < 0].sum())


def xxz_sparse_ground(n: int, delta: float) -> float:
    Sp = sp.csr_matrix([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.tocsr()
    Sz = sp.csr_matrix([[0.5, 0.0], [0.0, -0.5]])
    H = sp.csr_matrix((2**n, 2**n))
    for i in range(n - 1):
        left = sp.identity(2**i, format="csr")
        right = sp.identity(2 ** (n - i - 2), format="csr")
        for coef, A, B in ((-0.5, Sp, Sm), (-0.5, Sm, Sp), (-delta, Sz, Sz)):
            H = H + coef * sp.kron(sp.kron(left, sp.kron(A, B)), right)
    w = spla.eigsh(H.tocsr(), k=1, which="SA", tol=1e-14)[0]
    return float(w[0])


def main():
    fx = {
        "source": "reference test literals and notebook output (data), closed forms, brute force",
        "operators": harvest_operator_literals(),
        "ed": harvest_ed_literals(),
        "notebook_random_heisenberg_n6_h3_seed2021_evals": harvest_notebook_evals(),
        "closed_form": {
            "xx_chain_E0": {str(n): xx_closed_form(n) for n in (10, 16, 100, 200)},
            "ferro_xxz_delta1_E0": {str(n): -(n - 1) / 4 for n in (16, 100, 200)},
        },
        "brute_force": {
            "xxz_n10_delta0.5_E0": xxz_sparse_ground(10, 0.5),
            "xxz_n16_delta0.5_E0": xxz_sparse_ground(16, 0.5),
        },
        "random_fields": {
            "n6_h3_seed2021": np.random.RandomState(2021).uniform(-3.0, 3.0, size=6).tolist(),
            "n128_h0.5_seed2024": np.random.RandomState(2024).uniform(-0.5, 0.5, size=128).tolist(),
        },
    }
    OUT.write_text(json.dumps(fx, indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
