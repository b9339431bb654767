"""Writes a synthetic problem as TROVE checkpoint files, in the formats the
reference reads (test infrastructure):

  j0_extfield.chk            Fortran sequential records (TroveDipole.cpp:
                             18-73); records longer than `max_sub` bytes are
                             split into subrecords with negative leading
                             markers, as gfortran does beyond 2^31-9 bytes
                             (Util.cpp:89-109)
  j0eigen_quanta<J>.chk      text (TroveBasisSet.cpp:27-292)
  j0eigen_descr<J>_<g>.chk   text (TroveStates.cpp:15-131)
  j0eigen_vectors<J>_<g>.chk raw records (EigenVector.cpp:88)

The synthetic bases map onto the TROVE structure as: one vibrational class;
a case = one symmetry group of `gsize` consecutive vibrational functions in
one rotational level (K, tau); lambda = the member of the group; the J=0
function with the same class quanta (group q, member lambda) is v = q*gsize
+ lambda."""
import os
import struct

import numpy as np

from gain_engine import synth


def fortran_record(f, payload: bytes, max_sub: int = 2 ** 31 - 9):
    n = len(payload)
    chunks = [payload[i:i + max_sub] for i in range(0, n, max_sub)] or [b""]
    for i, ch in enumerate(chunks):
        first, last = i == 0, i == len(chunks) - 1
        f.write(struct.pack("<i", len(ch) if last else -len(ch)))
        f.write(ch)
        f.write(struct.pack("<i", len(ch) if first else -len(ch)))


def write_dipole(path, mu: np.ndarray, max_sub: int = 2 ** 31 - 9,
                 footer: bytes = b"End external field"):
    """mu in the piece layout [alpha][v_I][v_F]; the file holds, per
    component, the Fortran (column-major) matrix m(i, f) = mu[alpha][i][f]."""
    nc = mu.shape[1]
    with open(path, "wb") as f:
        fortran_record(f, b"Start external field", max_sub)
        fortran_record(f, struct.pack("<i", nc), max_sub)
        for a in range(3):
            fortran_record(f, struct.pack("<i", a + 1), max_sub)
            fortran_record(f, np.ascontiguousarray(mu[a].T).tobytes(),
                           max_sub)
        fortran_record(f, footer, max_sub)


def _rot_levels(J):
    """(K, tau) of each rotational level, in the basis' K-major order."""
    return [(0, J % 2)] + [(K, t) for K in range(1, J + 1) for t in (0, 1)]


def write_quanta(path, b: synth.Basis, gsize: int, reverse_cases=False):
    """`reverse_cases` lists the cases of every rotational level in reverse
    vibrational order (the file order then differs from the J=0 order, and
    the correlation has to match by content)."""
    lv = {kt: i for i, kt in enumerate(_rot_levels(b.J))}
    kval = b.kval()
    ncase = b.nprim // gsize
    order = list(range(ncase))
    if reverse_cases:
        per = b.nprim // len(_rot_levels(b.J)) // gsize
        order = [s * per + (per - 1 - j) for s in range(ncase // per)
                 for j in range(per)]
    L = ["Start Primitive basis set",
         "%8d%8d%8d%8d" % (ncase, gsize, b.nprim, 1)]
    for pos, c in enumerate(order):
        r0 = c * gsize
        il = lv[(int(kval[r0]), int(b.tau[r0]))]
        q = int(b.vib[r0]) // gsize
        L.append("%8d%6d%6d" % (gsize, il + 1, q + 1))
        for lam in range(gsize):
            L.append("%8d%6d%6d" % (pos * gsize + lam + 1, 1, lam + 1))
    levels = _rot_levels(b.J)
    L.append("%8d%8d" % (len(levels), 1))
    for i, (K, t) in enumerate(levels):
        L.append("%8d%8d%8d%8d%8d" % (i + 1, 1, b.J, K, t))
    L.append("End Primitive basis set")
    L.append("Start irreducible transformation")
    L.append("%8d" % b.matsize)
    L.append("".join("%8d" % n for n in b.ntot))
    for g in range(len(b.ntot)):
        L.extend("%8d" % n for n in b.N[g])
        nt, sd = b.ntot[g], b.sdeg[g]
        rep = b.repres[g].reshape(b.matsize, sd, nt) if nt else None
        for t in range(nt):
            for d in range(sd):
                L.append(" ".join("%24.16e" % rep[i, d, t]
                                  for i in range(b.matsize)))
    L.append("End irreducible transformation")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_block(directory, J, g, blk, rec_len, extra_high=(), label_of=None,
                shuffle_seed=None):
    """descr + vectors of one (J, irrep): states in energy order; record
    positions optionally shuffled; `extra_high` energies add states that the
    state filter drops (their records are written too); `label_of(s)` gives
    the level number printed for state s."""
    n = blk.n
    rng = np.random.default_rng(shuffle_seed)
    ntot = n + len(extra_high)
    order = rng.permutation(ntot) if shuffle_seed is not None \
        else np.arange(ntot)
    recs = np.zeros((ntot, rec_len))
    lines = ["Some TROVE header", "  (fingerprint lines)",
             "Start Quantum numbers and energies",
             "   J  Gamma  states", "   ----",
             "%8d%8d" % (ntot, rec_len)]
    for s in range(ntot):
        irec = int(order[s])
        if s < n:
            e = float(blk.energy[s])
            recs[irec] = blk.vector(s)
            lev = label_of(s) if label_of else s
        else:
            e = float(extra_high[s - n])
            recs[irec] = 0.0
            lev = n + (s - n) + 1000
        # 17 significant digits: the energies read back are the same doubles
        lines.append("%8d%8d%8d%8d %25.17e%8d%8d%8d" % (
            irec + 1, g + 1, lev + 1, 1, e, 0, 0, 0))
    lines.append("End Quantum numbers and energies")
    with open(os.path.join(directory, "j0eigen_descr%d_%d.chk" % (J, g + 1)),
              "w") as f:
        f.write("\n".join(lines) + "\n")
    recs.tofile(os.path.join(directory,
                             "j0eigen_vectors%d_%d.chk" % (J, g + 1)))


def write_problem(directory, prob: synth.Problem, scheme: str,
                  max_sub: int = 2 ** 31 - 9, extra_high=(), label_of=None,
                  shuffle_seed=None):
    gsize = synth._scheme(scheme)[2]
    write_dipole(os.path.join(directory, "j0_extfield.chk"), prob.mu, max_sub)
    write_quanta(os.path.join(directory, "j0eigen_quanta0.chk"),
                 synth.make_basis(0, prob.nc, scheme), gsize)
    for b in prob.bases:
        write_quanta(os.path.join(directory, "j0eigen_quanta%d.chk" % b.J),
                     b, gsize)
    for jind, b in enumerate(prob.bases):
        for g in range(prob.params.nrepres):
            blk = prob.blocks.get((jind, g))
            if blk is None:  # an irrep without states: an empty file pair
                blk = synth.Block(jind, g, np.zeros(0), np.zeros((0, 0)),
                                  b.ntot[g])
                write_block(directory, b.J, g, blk, b.ntot[g])
                continue
            write_block(directory, b.J, g, blk, blk.rec_len, extra_high,
                        label_of, shuffle_seed)
