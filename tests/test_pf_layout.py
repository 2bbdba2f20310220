"""CPU emulation of the factored form's Pfaffian step logic (k_fac_phi,
Pfaffian part) in its 2D block-cyclic register layout, lane by lane, against
an independent Pfaffian.

The kernel holds S(a + 8p, b + 8q) in slot (p, q), p <= q, of lane 8a + b
(diagonal slots: the whole 8 x 8 block), eliminates r = 2k, r + 1 per step with
partial pivoting on row r, exchanges the pivot column c with r + 1 by
overwriting row and column c with row r + 1, divides the pivot column once per
step and updates every slot of slot rows >= P = k / 4, leaving entries that left
the trailing matrix with values no one reads.  The emulation follows those steps
(the same LDS vectors R, A, C, D, started full of NaN) so that an indexing or
sign mistake in the design shows here; the GPU tests check the kernel itself
against the plain path and the oracle.  The reference Pfaffian is
Parlett-Reid with explicit row/column exchanges, itself checked against the
4 x 4 closed form and Pf^2 = det."""
import numpy as np
import pytest


def pf_ref(S):
    A = S.astype(complex).copy()
    n = A.shape[0]
    pf = 1.0 + 0j
    for k in range(0, n - 1, 2):
        kp = k + 1 + int(np.argmax(np.abs(A[k, k + 1:])))
        if kp != k + 1:
            A[[k + 1, kp], :] = A[[kp, k + 1], :]
            A[:, [k + 1, kp]] = A[:, [kp, k + 1]]
            pf = -pf
        if A[k, k + 1] == 0:
            return 0j
        pf *= A[k, k + 1]
        if k + 2 < n:
            tau = A[k, k + 2:] / A[k, k + 1]
            A[k + 2:, k + 2:] += np.outer(tau, A[k + 2:, k + 1]) - np.outer(A[k + 2:, k + 1], tau)
    return pf


def _key(v, j):
    # high word of |v|^2 with the low six bits replaced by 63 - j (ties to the smaller j)
    m = np.float64(abs(v) ** 2)
    hi = int(m.view(np.uint64)) >> 32
    return (hi & ~63) | (63 - j)


def emulate(S, plant=None):
    N2 = S.shape[0]
    H = N2 // 2
    T = (N2 + 7) // 8
    lane = np.arange(64)
    a, b = lane >> 3, lane & 7
    sv = {}
    for p in range(T):
        for q in range(p, T):
            i, j = a + 8 * p, b + 8 * q
            v = np.zeros(64, complex)
            ok = (i < N2) & (j < N2)
            v[ok] = S[i[ok], j[ok]]
            sv[(p, q)] = v
    L = 8 * T
    R, A, C, D = (np.full(L, np.nan + 0j) for _ in range(4))
    pf = 1.0 + 0j
    for P in range(T):
        for t in range(4):
            if 4 * P + t >= H - 1:
                break
            r = 8 * P + 2 * t
            best, c = -1, -1
            for q in range(P, T):
                for ln in np.nonzero(a == 2 * t)[0]:
                    j = b[ln] + 8 * q
                    if r < j < N2:
                        kk = _key(sv[(P, q)][ln], j)
                        if kk > best:
                            best, c = kk, j
            swap = c != r + 1
            Q, c8 = c >> 3, c & 7
            for ln in np.nonzero((a >> 1) == t)[0]:
                dst = A if a[ln] & 1 else R
                for q in range(P, T):
                    dst[b[ln] + 8 * q] = sv[(P, q)][ln]
            for ln in range(64):
                if b[ln] == c8:
                    for pp in range(P, Q + 1):
                        C[a[ln] + 8 * pp] = sv[(pp, Q)][ln]
                if a[ln] == c8:
                    for q in range(Q + 1, T):
                        D[b[ln] + 8 * q] = sv[(Q, q)][ln]
            s = R[c]
            if not abs(s) ** 2 > 1e-26:
                return 0j
            inv = 1 / s
            pf *= -s if (swap and plant != "exchange sign") else s
            if swap and plant != "no exchange":
                for ln in range(64):
                    if b[ln] == c8:
                        for pp in range(P, Q + 1):
                            sv[(pp, Q)][ln] = -A[a[ln] + 8 * pp]
                    if a[ln] == c8:
                        for q in range(Q, T):
                            sv[(Q, q)][ln] = A[b[ln] + 8 * q]
            newC, newR = C.copy(), R.copy()  # one instruction stream: all lanes read before any writes
            for x in range(r + 2, N2):
                xc = swap and x == c
                low = x < 8 * Q + 8
                v = C[r + 1] if xc else (C[x] if low else D[x])
                newC[x] = v * (inv if (xc or low) else -inv)
                if xc:
                    newR[c] = R[r + 1]
            C[:], R[:] = newC, newR
            for pp in range(P, T):
                Ri, Ci = R[a + 8 * pp], C[a + 8 * pp]
                for q in range(pp, T):
                    sv[(pp, q)] = sv[(pp, q)] - Ci * R[b + 8 * q] + Ri * C[b + 8 * q]
    i0, j0 = N2 - 2, N2 - 1
    return pf * sv[(i0 >> 3, j0 >> 3)][(i0 & 7) * 8 + (j0 & 7)]


def test_reference_pfaffian_pinned():
    rng = np.random.default_rng(11)
    S = rng.normal(size=(4, 4))
    S = S - S.T
    assert abs(pf_ref(S) - (S[0, 1] * S[2, 3] - S[0, 2] * S[1, 3] + S[0, 3] * S[1, 2])) < 1e-12
    X = rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12))
    S = X - X.T
    assert abs(pf_ref(S) ** 2 - np.linalg.det(S)) < 1e-9 * abs(np.linalg.det(S))


@pytest.mark.parametrize("n2", [4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44])
def test_layout_matches_reference(n2):
    rng = np.random.default_rng(100 + n2)
    X = rng.normal(size=(n2, n2)) + 1j * rng.normal(size=(n2, n2))
    S = X - X.T
    assert abs(emulate(S) - pf_ref(S)) <= 1e-12 * abs(pf_ref(S))
    # the overlap matrix of the factored form: [[e^{2i phi} M0, -1], [1, -Z0]]
    h = n2 // 2
    M = rng.normal(size=(h, h)) + 1j * rng.normal(size=(h, h))
    Z = rng.normal(size=(h, h))
    S = np.block([[np.exp(0.6j) * (M - M.T), -np.eye(h)], [np.eye(h), -(Z - Z.T)]])
    assert abs(emulate(S) - pf_ref(S)) <= 1e-12 * abs(pf_ref(S))


@pytest.mark.parametrize("plant", ["exchange sign", "no exchange"])
def test_planted_mistakes_are_caught(plant):
    rng = np.random.default_rng(7)
    X = rng.normal(size=(24, 24)) + 1j * rng.normal(size=(24, 24))
    S = X - X.T
    assert abs(emulate(S, plant) - pf_ref(S)) > 1e-6 * abs(pf_ref(S))


@pytest.mark.parametrize("h", [2, 4, 10, 20, 22])
def test_schur_complement_pfaffian_identity(h):
    """The factored form's opt-in Pfaffian route: for the species matrix
    S = [[A, -1], [1, B]] (A = e^{2i phi} M0 complex skew, B = -Z_s real skew
    and invertible), Pf(S) = Pf(B) Pf(A + B^-1), with the cscale factor split
    as Pf(c S) = Pf(c B) Pf(c (A + B^-1)); the h x h Pfaffian through the
    kernel's 2D layout emulation.  A dropped sign (A - B^-1) is caught."""
    rng = np.random.default_rng(300 + h)
    X = rng.normal(size=(h, h)) + 1j * rng.normal(size=(h, h))
    A = np.exp(0.8j) * (X - X.T)
    Y = rng.normal(size=(h, h))
    B = Y - Y.T
    S = np.block([[A, -np.eye(h)], [np.eye(h), B]])
    full = pf_ref(S)
    Binv = np.linalg.inv(B)
    assert abs(pf_ref(B) * pf_ref(A + Binv) - full) <= 1e-10 * abs(full)
    c = 1.7
    assert abs(pf_ref(c * B) * emulate(c * (A + Binv)) - pf_ref(c * S)) <= 1e-10 * abs(pf_ref(c * S))
    assert abs(pf_ref(B) * pf_ref(A - Binv) - full) > 1e-6 * abs(full)
