"""Writes tests/golden/hubbard12.FCIDUMP: the open Hubbard chain of the reference's dmrg/tests/HUBBARD12 input
(12 sites, 12 electrons, U = 1 on (ii|ii), t = +1 on h(i+1, i)), in the FCIDUMP format read by
mpsxx_amd.qc.read_fcidump.  Run from the repository root: python tests/golden/make_hubbard12_fcidump.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mpsxx_amd.qc import write_fcidump  # noqa: E402

L = 12
h = np.zeros((L, L))
for i in range(L - 1):
    h[i, i + 1] = h[i + 1, i] = 1.0
v = np.zeros((L, L, L, L))
for i in range(L):
    v[i, i, i, i] = 1.0
write_fcidump(os.path.join(ROOT, "tests", "golden", "hubbard12.FCIDUMP"), h, v, 0.0, nelec=L, ms2=0)
