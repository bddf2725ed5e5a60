"""Writes tests/golden/c1_superlu.npz: the C1 (2D dam break, h = 0.05) PSPG system assembled
by the oracle at the hydrostatic initial state, solved by SciPy's SuperLU with COLAMD column
ordering — the published algorithm Eigen's SparseLU<..., COLAMDOrdering<int>> is built on
(the reference's solver, MomContEquation.hpp:6-12).  Run once in a container with SciPy;
the fixture is data (row pointer, b norm, solution).
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pfem_amd import meshgen as mg  # noqa: E402


def main():
    m = mg.dambreak_2d()
    q = mg.pspg_hydrostatic(m)
    S = O.pspg_assemble(m, q, dt=1e-3)
    A = S.to_scipy().tocsc()
    lu = spla.splu(A, permc_spec="COLAMD")
    x = lu.solve(S.b)
    rr = np.linalg.norm(S.b - A @ x) / np.linalg.norm(S.b)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_superlu.npz")
    np.savez_compressed(out, n=S.n, nnz=len(S.vals), row_ptr=S.row_ptr, b_norm=np.linalg.norm(S.b), x=x, relres=rr)
    print(f"wrote {out}: n={S.n} nnz={len(S.vals)} relres={rr:.3e}")


if __name__ == "__main__":
    main()
