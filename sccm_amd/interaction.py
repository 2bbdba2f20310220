"""Two-body interaction in the m-scheme list form PNAMP consumes.

The energy kernel (``Codes/PNAMP/PNAMP_Energy_OpenMP.f90:31-80``) walks a list
of antisymmetrised matrix elements V_abcd = <ab|V|cd> (a, b, c, d are 1-based
m-state indices).  The list holds exactly the entries with a < b, c < d,
a + b <= c + d that conserve the total isospin projection, M and parity, in
nested (a, b, c, d) order.  That rule reproduces the record counts of the
reference's binary lists (``Codes/int/*/IMSRG_*_m1m2m3m4.dat``: 14442 for pf,
10330 for pf5g9) and is the only one under which the eight-term sum of
``HFB_Energy_COMPLEX`` equals sum_{a<b,c<d} V_abcd <a+ b+ d c>.

Synthetic interactions are built the way the reference's ``Hamiltonian``
uncouples J-scheme matrix elements (``Codes/GCM/HFBA_Hamiltonian.f90:295-371``):
a normalised two-body state |alpha beta; J M> is expanded in m-scheme pairs with
Clebsch-Gordan coefficients, which makes the list rotation invariant by
construction.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from math import factorial, sqrt

import numpy as np

from .basis import Basis


@lru_cache(maxsize=None)
def cg(j1: int, m1: int, j2: int, m2: int, j: int, m: int) -> float:
    """Clebsch-Gordan <j1 m1 j2 m2 | j m>, all arguments doubled (Racah formula,
    Condon-Shortley phase; same convention as ``CG`` in
    ``Codes/GCM/VAPHFB_ThreeJ.f90:51-151``)."""
    if m1 + m2 != m or abs(m1) > j1 or abs(m2) > j2 or abs(m) > j:
        return 0.0
    if (j1 + m1) % 2 or (j2 + m2) % 2 or (j + m) % 2:
        return 0.0
    if j1 + j2 < j or abs(j1 - j2) > j or (j1 + j2 + j) % 2:
        return 0.0
    a = (j1 + j2 - j) // 2
    b = (j1 - j2 + j) // 2
    c = (-j1 + j2 + j) // 2
    pre = sqrt((j + 1) * factorial(a) * factorial(b) * factorial(c)
               / factorial((j1 + j2 + j) // 2 + 1))
    pre *= sqrt(factorial((j1 + m1) // 2) * factorial((j1 - m1) // 2)
                * factorial((j2 + m2) // 2) * factorial((j2 - m2) // 2)
                * factorial((j + m) // 2) * factorial((j - m) // 2))
    s = 0.0
    for k in range(0, a + 1):
        d1 = a - k
        d2 = (j1 - m1) // 2 - k
        d3 = (j2 + m2) // 2 - k
        d4 = (j - j2 + m1) // 2 + k
        d5 = (j - j1 - m2) // 2 + k
        if min(d1, d2, d3, d4, d5) < 0:
            continue
        s += (-1) ** k / (factorial(k) * factorial(d1) * factorial(d2)
                          * factorial(d3) * factorial(d4) * factorial(d5))
    return pre * s


@dataclass
class Interaction:
    """m-scheme list + one-body part, as held in the reference's ``H`` module
    variables (``H%ka..kd, H%ME2BM, H%zME1BM, H%E0``)."""

    ka: np.ndarray   # int32, 1-based
    kb: np.ndarray
    kc: np.ndarray
    kd: np.ndarray
    me: np.ndarray   # float64
    t1b: np.ndarray  # (n, n) float64, t1b[a, c] = H%zME1BM(a, c)
    e0: float = 0.0

    @property
    def nme(self) -> int:
        return int(self.me.size)


def list_indices(b: Basis) -> np.ndarray:
    """(L, 4) 1-based index quadruples in the reference's nested order."""
    n = b.n
    t, tm, l = b.t.astype(int), b.twom.astype(int), b.l.astype(int)
    out = []
    for a in range(n):
        for bb in range(a + 1, n):
            for c in range(n):
                for d in range(c + 1, n):
                    if a + bb > c + d:
                        continue
                    if t[a] + t[bb] != t[c] + t[d]:
                        continue
                    if tm[a] + tm[bb] != tm[c] + tm[d]:
                        continue
                    if (l[a] + l[bb] - l[c] - l[d]) % 2:
                        continue
                    out.append((a + 1, bb + 1, c + 1, d + 1))
    return np.asarray(out, dtype=np.int32).reshape(-1, 4)


def _global_orbit(b: Basis) -> np.ndarray:
    """Orbit label that includes the species: protons 0..norb-1, neutrons
    norb..2 norb-1.  Because protons come first and orbits are contiguous,
    a < b implies go[a] <= go[b]."""
    norb = len(b.orbits)
    return b.orbit + np.where(b.t > 0, norb, 0)


def _pair_coeff(b: Basis, go, a: int, c: int, J2: int) -> float:
    """<a c | (alpha gamma) J M> for 0-based m-states a < c, normalised
    coupled state, antisymmetric m-scheme pair a+_a a+_c |0>."""
    ja, jc = int(b.twoj[a]), int(b.twoj[c])
    ma, mc = int(b.twom[a]), int(b.twom[c])
    M2 = ma + mc
    if go[a] != go[c]:
        return cg(ja, ma, jc, mc, J2, M2)
    return (cg(ja, ma, jc, mc, J2, M2) - cg(ja, mc, jc, ma, J2, M2)) / sqrt(2.0)


def random_tbme_channels(b: Basis, sigma: float, rng: np.random.Generator) -> dict:
    """Seeded normalised J-scheme matrix elements V_J(alpha beta, gamma delta).

    Channels are (J, Tz, parity); inside a channel the matrix over orbit pairs
    alpha <= beta is real symmetric with N(0, sigma) entries.  Identical-orbit
    pairs only carry even J (antisymmetry).  Returns {(J2, go_a, go_b, go_c, go_d): V}.
    """
    norb = len(b.orbits)
    orbs = []  # global orbit list: (species, n, l, twoj)
    for sp in (-1, 1):
        for (n, l, tj) in b.orbits:
            orbs.append((sp, n, l, tj))
    pairs = [(i, j) for i in range(2 * norb) for j in range(i, 2 * norb)]
    chans: dict = {}
    for (i, j) in pairs:
        ji, jj = orbs[i][3], orbs[j][3]
        for J2 in range(abs(ji - jj), ji + jj + 1, 2):
            if i == j and (J2 // 2) % 2:
                continue
            key = (J2, orbs[i][0] + orbs[j][0], (orbs[i][2] + orbs[j][2]) % 2)
            chans.setdefault(key, []).append((i, j))
    table = {}
    for key in sorted(chans):
        J2 = key[0]
        members = chans[key]
        m = len(members)
        mat = rng.normal(0.0, sigma, size=(m, m))
        mat = np.triu(mat) + np.triu(mat, 1).T
        for p in range(m):
            for q in range(m):
                table[(J2,) + members[p] + members[q]] = float(mat[p, q])
    return table


def uncouple(b: Basis, table: dict) -> np.ndarray:
    """m-scheme values for every list entry from J-scheme channels."""
    idx = list_indices(b) - 1
    go = _global_orbit(b)
    me = np.zeros(len(idx))
    for e, (a, bb, c, d) in enumerate(idx):
        ja, jb, jc, jd = (int(b.twoj[x]) for x in (a, bb, c, d))
        M2 = int(b.twom[a] + b.twom[bb])
        lo = max(abs(ja - jb), abs(jc - jd), abs(M2))
        hi = min(ja + jb, jc + jd)
        s = 0.0
        for J2 in range(lo, hi + 1, 2):
            key = (J2, int(go[a]), int(go[bb]), int(go[c]), int(go[d]))
            v = table.get(key)
            if v is None:
                continue
            s += _pair_coeff(b, go, a, bb, J2) * _pair_coeff(b, go, c, d, J2) * v
        me[e] = s
    return me


def synthetic_interaction(b: Basis, tbme_sigma: float, seed_tbme: int,
                          seed_spe: int, spe_range=(-5.0, 0.0)) -> Interaction:
    """Seeded synthetic shell-model interaction: J-coupled TBMEs N(0, sigma),
    single-particle energies uniform in ``spe_range`` per (orbit, species)."""
    rng = np.random.default_rng(seed_tbme)
    table = random_tbme_channels(b, tbme_sigma, rng)
    idx = list_indices(b)
    me = uncouple(b, table)
    rs = np.random.default_rng(seed_spe)
    norb = len(b.orbits)
    spe = rs.uniform(spe_range[0], spe_range[1], size=2 * norb)
    go = _global_orbit(b)
    t1b = np.diag(spe[go]).astype(np.float64)
    return Interaction(idx[:, 0].copy(), idx[:, 1].copy(), idx[:, 2].copy(),
                       idx[:, 3].copy(), me, t1b, 0.0)


def parse_jscheme_tbme(text: str, b: Basis) -> dict:
    """Shell-model J-scheme file ``t1..t4, n1..n4, lj1..lj4, J, V`` (comma or
    blank separated; t = 1 proton, 0 neutron; lj = (2l + 2j - 1)/2), as read by
    ``Codes/PNAMP/IMSRG_GenerateH_JTMT.f90:108``, into the channel table of
    ``uncouple``.  The files store identical-orbit elements unnormalised:
    V_norm = V / sqrt((1 + d_ab)(1 + d_cd)).  Pairs are put in ascending
    orbit order (protons first) with the exchange phase -(-1)^(ja + jb - J)."""
    norb = len(b.orbits)
    lj_of = [(2 * l + tj - 1) // 2 for (_, l, tj) in b.orbits]
    tj_of = [tj for (_, _, tj) in b.orbits]

    def go(t, lj):
        return lj_of.index(lj) + (0 if t == 1 else norb)

    table = {}
    for line in text.splitlines():
        f = line.replace(",", " ").split()
        if len(f) < 14:
            continue
        t1, t2, t3, t4, _, _, _, _, l1, l2, l3, l4, J = (int(x) for x in f[:13])
        V = float(f[13])
        a, bb, c, d = go(t1, l1), go(t2, l2), go(t3, l3), go(t4, l4)
        ja, jb, jc, jd = (tj_of[x % norb] for x in (a, bb, c, d))
        v = V / sqrt((1 + (a == bb)) * (1 + (c == d)))
        if a > bb:
            a, bb = bb, a
            v *= -(-1) ** ((ja + jb) // 2 - J)
        if c > d:
            c, d = d, c
            v *= -(-1) ** ((jc + jd) // 2 - J)
        table[(2 * J, a, bb, c, d)] = v
        table[(2 * J, c, d, a, bb)] = v
    return table


def read_mscheme_binary(path: str) -> np.ndarray:
    """Fortran sequential unformatted list, one real*8 per record
    (``HFBA_Hamiltonian.f90:373-382``)."""
    raw = np.fromfile(path, dtype=np.dtype([("h", "<i4"), ("v", "<f8"), ("t", "<i4")]))
    if raw.size and not (np.all(raw["h"] == 8) and np.all(raw["t"] == 8)):
        raise ValueError(f"{path}: not a list of 8-byte records")
    return raw["v"].astype(np.float64)


def read_spe(path: str, b: Basis) -> tuple[np.ndarray, float]:
    with open(path) as f:
        return parse_spe(f.read(), b)


def parse_spe(text: str, b: Basis) -> tuple[np.ndarray, float]:
    """One-body file ``t LJ n1 n2 e`` with E0 on the first line, mapped to the
    m-scheme as ``kinetic`` does (``Codes/GCM/HFBA_Hamiltonian1B.f90:28-63``):
    t=0 neutron, t=1 proton; non-zero only between equal (t, lj, m)."""
    lines = [ln.split() for ln in text.splitlines() if ln.strip()]
    e0 = float(lines[0][0])
    ee = {}
    for ln in lines[1:]:
        t1, lj, n1, n2, e = int(ln[0]), int(ln[1]), int(ln[2]), int(ln[3]), float(ln[4])
        sp = 1 if t1 == 0 else -1
        ee[(sp, n1 + 1, n2 + 1, lj)] = e
    n = b.n
    t1b = np.zeros((n, n))
    lj = b.lj
    for k in range(n):
        for l_ in range(n):
            if b.t[k] != b.t[l_] or lj[k] != lj[l_] or b.twom[k] != b.twom[l_]:
                continue
            t1b[l_, k] = ee.get((int(b.t[k]), int(b.nq[l_]), int(b.nq[k]), int(lj[k])), 0.0)
    return t1b, e0
