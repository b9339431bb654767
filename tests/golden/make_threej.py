"""Writes threej_jmax31.json: exact Wigner 3j values (sympy, third party) for
every entry of the reference's 3j table (GpuManager.cpp:164-184) with
jmax = 31: (J_I 1 J_F; K_I, K_F-K_I, -K_F) for J_F in J_I-1..J_I+1 and
K_F in K_I-1..K_I+1.  Run: python tests/golden/make_threej.py"""
import json
import os

from sympy.physics.wigner import wigner_3j

JMAX = 31
out = []
for jI in range(JMAX + 1):
    for jF in range(max(jI - 1, 0), min(jI + 1, JMAX) + 1):
        for kI in range(jI + 1):
            for kF in range(max(kI - 1, 0), min(kI + 1, jF) + 1):
                v = wigner_3j(jI, 1, jF, kI, kF - kI, -kF)
                out.append([jI, kI, jF, kF, repr(float(v.evalf(30)))])
path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                    "threej_jmax31.json")
with open(path, "w") as f:
    json.dump({"jmax": JMAX, "source": "sympy.physics.wigner.wigner_3j",
               "entries": out}, f, separators=(",", ":"))
print(len(out), "entries ->", path)
