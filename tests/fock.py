"""Exact Fock-space reference for tiny model spaces (n <= 12 m-states).

Independent of the oracle's formulas: states are built as vectors in the
2^n-dimensional Fock space, operators as sparse matrices, and every quantity
the PNAMP path produces is computed by brute force:

* |Phi> = exp(1/2 sum Z_ab c+_a c+_b)|0> / norm with Z = V U^{-1}
  (Thouless; beta_k = sum_l U_lk c_l + V_lk c+_l annihilates it);
* R(alpha, beta, gamma) = exp(-i alpha Jz) exp(-i beta Jy) exp(-i gamma Jz)
  and the gauge rotations exp(i phi_p Z) exp(i phi_n N) applied to vectors;
* transition densities <Phi0| c+_j c_i O |Phi1> / <Phi0|O|Phi1>;
* the Hamiltonian sum_ac t_ac c+_a c_c + sum_list V_abcd (c+_a c+_b c_d c_c + partner);
* angular-momentum projectors P^J_{KK'} = sum_alpha |J K alpha><J K' alpha|
  from the eigenvectors of J^2 and Condon-Shortley ladder operators.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply


class Fock:
    def __init__(self, basis):
        self.b = basis
        n = basis.n
        if n > 14:
            raise ValueError("Fock space too large")
        self.n = n
        self.dim = 1 << n
        self.states = np.arange(self.dim, dtype=np.int64)
        self._cd = [self._creation(k) for k in range(n)]
        self._c = [m.T.tocsr() for m in self._cd]
        self.num = np.array([bin(s).count("1") for s in range(self.dim)])
        pmask = sum(1 << k for k in range(n) if basis.t[k] < 0)
        self.nprot = np.array([bin(s & pmask).count("1") for s in range(self.dim)])
        self.nneut = self.num - self.nprot
        self.mtot2 = np.array([sum(int(basis.twom[k]) for k in range(n) if (s >> k) & 1)
                               for s in range(self.dim)])
        self.Jy = self.one_body(self._jmat("y"))
        self.Jp = self.one_body(self._jmat("+"))
        self.Jm = self.one_body(self._jmat("-"))

    def _creation(self, k):
        rows, cols, vals = [], [], []
        for s in range(self.dim):
            if (s >> k) & 1:
                continue
            sign = -1.0 if bin(s & ((1 << k) - 1)).count("1") % 2 else 1.0
            rows.append(s | (1 << k))
            cols.append(s)
            vals.append(sign)
        return sp.csr_matrix((vals, (rows, cols)), shape=(self.dim, self.dim))

    def cd(self, k):
        return self._cd[k]

    def c(self, k):
        return self._c[k]

    def one_body(self, O):
        """sum_ab O_ab c+_a c_b as a sparse many-body matrix."""
        M = sp.csr_matrix((self.dim, self.dim), dtype=complex)
        for a in range(self.n):
            for b in range(self.n):
                if O[a, b] != 0:
                    M = M + O[a, b] * (self.cd(a) @ self.c(b))
        return M.tocsr()

    def _jmat(self, comp):
        b, n = self.b, self.n
        O = np.zeros((n, n), dtype=complex)
        for a in range(n):
            for c in range(n):
                if b.t[a] != b.t[c] or b.orbit[a] != b.orbit[c]:
                    continue
                j = b.twoj[a] / 2.0
                ma, mc = b.twom[a] / 2.0, b.twom[c] / 2.0
                if comp in ("+", "y") and abs(ma - (mc + 1)) < 1e-9:
                    jp = np.sqrt(j * (j + 1) - mc * (mc + 1))
                    O[a, c] += jp if comp == "+" else jp / (2.0 * 1j)
                if comp in ("-", "y") and abs(ma - (mc - 1)) < 1e-9:
                    jm = np.sqrt(j * (j + 1) - mc * (mc - 1))
                    O[a, c] += jm if comp == "-" else -jm / (2.0 * 1j)
        return O

    # ---------------------------------------------------------------- states
    def hfb_state(self, U, V):
        Z = V @ np.linalg.inv(U)
        P = sp.csr_matrix((self.dim, self.dim), dtype=complex)
        for a in range(self.n):
            for b in range(a + 1, self.n):
                if Z[a, b] != 0:
                    P = P + Z[a, b] * (self.cd(a) @ self.cd(b))
        v = np.zeros(self.dim, dtype=complex)
        v[0] = 1.0
        term = v.copy()
        for k in range(1, self.n // 2 + 1):
            term = P @ term / k
            v = v + term
        return v / np.linalg.norm(v)

    # ---------------------------------------------------------------- operators
    def rotate(self, v, alp, bet, gam, phip=0.0, phin=0.0):
        """R(alp,bet,gam) exp(i phip Z) exp(i phin N) v."""
        w = v * np.exp(1j * (phip * self.nprot + phin * self.nneut))
        w = w * np.exp(-1j * gam * self.mtot2 / 2.0)
        if bet != 0.0:
            w = expm_multiply(-1j * bet * self.Jy, w)
        return w * np.exp(-1j * alp * self.mtot2 / 2.0)

    def hamiltonian(self, inter):
        n = self.n
        H = self.one_body(inter.t1b.astype(complex))
        for q in range(inter.nme):
            a, b_, c, d = int(inter.ka[q]) - 1, int(inter.kb[q]) - 1, int(inter.kc[q]) - 1, int(inter.kd[q]) - 1
            v = inter.me[q]
            if v == 0:
                continue
            H = H + v * (self.cd(a) @ self.cd(b_) @ self.c(d) @ self.c(c))
            if a + b_ != c + d:
                H = H + v * (self.cd(c) @ self.cd(d) @ self.c(b_) @ self.c(a))
        return H.tocsr()

    def densities(self, v0, w1):
        """rho_ij = <v0|c+_j c_i|w1>/<v0|w1>, kappa10_ij = <v0|c_j c_i|w1>/<v0|w1>,
        kappa01_ij = <v0|c+_i c+_j|w1>/<v0|w1>."""
        n = self.n
        ov = np.vdot(v0, w1)
        rho = np.zeros((n, n), dtype=complex)
        k10 = np.zeros((n, n), dtype=complex)
        k01 = np.zeros((n, n), dtype=complex)
        cw = [self.c(i) @ w1 for i in range(n)]
        for i in range(n):
            for j in range(n):
                rho[i, j] = np.vdot(v0, self.cd(j) @ cw[i]) / ov
                k10[i, j] = np.vdot(v0, self.c(j) @ cw[i]) / ov
                k01[i, j] = np.vdot(v0, self.cd(i) @ (self.cd(j) @ w1)) / ov
        return rho, k10, k01

    # ---------------------------------------------------------------- projection
    def projectors(self, zval, nval, jmax):
        """{(J, K1, K2): sparse-free dense operator restricted to the (Z, N)
        sector} as (sector index array, dict of matrices)."""
        sec = np.nonzero((self.nprot == zval) & (self.nneut == nval))[0]
        pos = {s: i for i, s in enumerate(sec)}
        Jp = self.Jp[sec][:, sec].toarray()
        Jm = self.Jm[sec][:, sec].toarray()
        m2 = self.mtot2[sec]
        J2op = Jm @ Jp + np.diag((m2 / 2.0) ** 2 + m2 / 2.0)
        P = {}
        for J in range(0, jmax + 1):
            idx = np.nonzero(m2 == 2 * J)[0]
            if idx.size == 0:
                continue
            sub = J2op[np.ix_(idx, idx)]
            w, vec = np.linalg.eigh((sub + sub.conj().T) / 2)
            hw = vec[:, np.abs(w - J * (J + 1)) < 1e-8]
            if hw.shape[1] == 0:
                continue
            full = np.zeros((sec.size, hw.shape[1]), dtype=complex)
            full[idx, :] = hw
            kets = {J: full}
            cur = full
            for K in range(J, -J, -1):
                cur = Jm @ cur / np.sqrt((J + K) * (J - K + 1))
                kets[K - 1] = cur
            for K1 in range(-J, J + 1):
                for K2 in range(-J, J + 1):
                    P[(J, K1, K2)] = kets[K1] @ kets[K2].conj().T
        return sec, P
