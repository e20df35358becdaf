"""CPU restatement of the reference CCD/(T) amplitude iteration.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: tests/, the smoke check and
the CPU-baseline leg of bench.py may import it as the *checker*.  The product
(the HIP library and its hosts) never imports, links or calls it.

It follows the reference's ordered-pair algorithm step by step (file:line below
are relative to deltafull-CCMkspace/ of the reference):

  basis          ccm_main.f90:878-1197 (allocate_sp_data, setup_sp_data)
  channels       ccm_iter.f90:175-811 (setup_channel_structures),
                 ccm_iter.f90:1816-2212 (setup_ph_channel_structures)
  Fock, E0       ccm_main.f90:1586-1684 (normal_ordered_hamiltonian)
  MBPT2 start    ccm_iter.f90:2329-2405 (setup_t_amplitudes)
  residual       ccm_t2_eqn.f90:4-1304 (t2_intermediate), 1312-1518 (I6/I7)
  update + DIIS  ccm_diis.f90:74-293 (diis_t1t2, diis_setup)
  energy         ccm_energy.f90:4-71 (ccd_energy_save)
  driver loop    ccm_iter.f90:1-170 (ccd_iter)
  Anderson       the engine's own accelerator (no reference counterpart),
                 restated in solve_anderson so the GPU trajectory has a checker
  (T)            ccm_triples.f90:174-294, 528-735; ccm_t3full_channels.f90:4-284
  Minnesota      chiral_module_andreas_with_delta.f90:8980-9111 (vmom_minnesota)

Pins (tests/test_oracle_*.py): per-iteration CCD energies, E0, Fock and (T)
against fixtures written by the compiled reference (tests/golden/), closed forms
for the basis, brute-force pair enumeration for the channels, a dense
4-index residual on a tiny lattice, and a DIIS toy problem with a known answer.
"""
from __future__ import annotations

import numpy as np

# Constants exactly as the reference declares them (ccm_modules.f90:91-95).  The
# nucleon mass is a default-real literal, so it is the float32 value of 939.565.
HBARC = 197.326968
P_MASS = float(np.float32(939.565))
PI = 3.14159265358979


def reference_index_order(arr):
    """Ascending order of arr as the reference's `index` leaves it
    (ccm_main.f90:2654-2732): quicksort, median of three as pivot, partitions
    below 7 elements finished by insertion, the larger part deferred on a
    stack.  Equal keys come out in that routine's order, which decides which of
    several degenerate momenta become holes.  Returns 0-based positions."""
    n = len(arr)
    ix = list(range(n + 1))                 # 1-based: ix[r] = element at rank r
    key = lambda r: arr[ix[r] - 1]
    stack = []
    lo, hi = 1, n
    while True:
        if hi - lo < 7:
            for j in range(lo + 1, hi + 1):
                moving = ix[j]
                v = arr[moving - 1]
                i = j - 1
                while i >= 1 and key(i) > v:
                    ix[i + 1] = ix[i]
                    i -= 1
                ix[i + 1] = moving
            if not stack:
                return [x - 1 for x in ix[1:]]
            hi = stack.pop()
            lo = stack.pop()
            continue
        mid = (lo + hi) // 2
        ix[mid], ix[lo + 1] = ix[lo + 1], ix[mid]
        if key(lo + 1) > key(hi):
            ix[lo + 1], ix[hi] = ix[hi], ix[lo + 1]
        if key(lo) > key(hi):
            ix[lo], ix[hi] = ix[hi], ix[lo]
        if key(lo + 1) > key(lo):
            ix[lo + 1], ix[lo] = ix[lo], ix[lo + 1]
        pivot = ix[lo]
        v = arr[pivot - 1]
        i, j = lo + 1, hi
        while True:
            i += 1
            while key(i) < v:
                i += 1
            j -= 1
            while key(j) > v:
                j -= 1
            if j < i:
                break
            ix[i], ix[j] = ix[j], ix[i]
        ix[lo] = ix[j]
        ix[j] = pivot
        if hi - i + 1 >= j - lo:
            stack += [i, hi]
            hi = j - 1
        else:
            stack += [lo, j - 1]
            lo = i


class Lattice:
    """Single-particle plane-wave basis on the periodic cubic lattice.

    Orbital arrays follow setup_sp_data (ccm_main.f90:1003-1197): momenta
    k = (2 pi n + theta)/L per axis, e = hbar^2 k^2 / 2m, states ordered by k^2
    and, within one momentum, sz fastest then tz; the first A states are holes.
    Mesh points are sorted by k^2 with the reference's own quicksort
    (reference_index_order), so degenerate momenta come out in its order and
    orbital indices, the hole set of an open shell included, equal the
    reference's.
    """

    def __init__(self, nx, ny, nz, sz, tz, kx, ky, kz, e, nholes, box, volume,
                 kf=None, rho=None, nmax=None):
        self.n = np.stack([np.asarray(nx), np.asarray(ny), np.asarray(nz)], 1).astype(np.int64)
        self.sz = np.asarray(sz, dtype=np.int64)
        self.tz = np.asarray(tz, dtype=np.int64)
        self.k = np.stack([np.asarray(kx), np.asarray(ky), np.asarray(kz)], 1).astype(np.float64)
        self.e = np.asarray(e, dtype=np.float64)
        self.nsp = len(self.e)
        self.A = int(nholes)
        self.box = float(box)
        self.volume = float(volume)
        self.kf = kf
        self.rho = rho
        self.nmax = nmax if nmax is not None else int(np.abs(self.n).max())

    @classmethod
    def build(cls, A, matter, rho, nmax, twist=(0.0, 0.0, 0.0)):
        """allocate_sp_data + setup_sp_data (ccm_main.f90:924-951, 1003-1164)."""
        if matter == "pnm":
            tzs = [1]
            kf = (3.0 * PI ** 2 * rho) ** (1.0 / 3.0)
        elif matter == "snm":
            tzs = [-1, 1]
            kf = (3.0 * PI ** 2 * rho / 2.0) ** (1.0 / 3.0)
        else:
            raise ValueError(f"unknown matter {matter!r}")
        L = (float(A) / rho) ** (1.0 / 3.0)
        # mesh order of the reference: 0, -1, +1, -2, +2, ...
        ns = [0]
        for i in range(1, nmax + 1):
            ns += [-i, i]
        th = [float(t) for t in twist]
        mesh = []
        for nxv in ns:
            for nyv in ns:
                for nzv in ns:
                    kx = (2.0 * nxv * PI + th[0]) / L
                    ky = (2.0 * nyv * PI + th[1]) / L
                    kz = (2.0 * nzv * PI + th[2]) / L
                    mesh.append((kx * kx + ky * ky + kz * kz, nxv, nyv, nzv, kx, ky, kz))
        order = reference_index_order([m[0] for m in mesh])
        cols = {k: [] for k in ("nx", "ny", "nz", "sz", "tz", "kx", "ky", "kz", "e")}
        for m in order:
            k2, nxv, nyv, nzv, kx, ky, kz = mesh[m]
            # within a momentum: (sz=-1,tz_lo), (sz=+1,tz_lo), (sz=-1,tz_hi), ...
            for tzv in tzs:
                for szv in (-1, 1):
                    cols["nx"].append(nxv); cols["ny"].append(nyv); cols["nz"].append(nzv)
                    cols["sz"].append(szv); cols["tz"].append(tzv)
                    cols["kx"].append(kx); cols["ky"].append(ky); cols["kz"].append(kz)
                    cols["e"].append((kx ** 2 + ky ** 2 + kz ** 2) * HBARC ** 2 / P_MASS / 2.0)
        return cls(cols["nx"], cols["ny"], cols["nz"], cols["sz"], cols["tz"],
                   cols["kx"], cols["ky"], cols["kz"], cols["e"], A, L, L ** 3,
                   kf=kf, rho=rho, nmax=nmax)


# --------------------------------------------------------------------------- #
# interactions: vectorized antisymmetrized elements <pq||rs>
# --------------------------------------------------------------------------- #

def _sigma_dot_sigma(m1, m2, m3, m4):
    """<m1 m2| sigma1.sigma2 |m3 m4> for m = +-1 (chp_sigma_dot_sigma,
    chiral_module_andreas_with_delta.f90:2975-3031): diagonal m1*m2, spin
    exchange 2, everything else 0."""
    diag = (m1 == m3) & (m2 == m4)
    exch = (m1 == m4) & (m2 == m3) & (m1 != m2)
    return np.where(diag, (m1 * m2).astype(np.float64), 0.0) + np.where(exch, 2.0, 0.0)


def minnesota_direct(lat: Lattice, p, q, r, s):
    """Direct Minnesota element v(p,q,r,s) (vmom_minnesota,
    chiral_module_andreas_with_delta.f90:8980-9111), real, in MeV.
    Constants are the reference's default-real literals (float32 values)."""
    f32 = lambda x: float(np.float32(x))
    v0r, v0t, v0s = 200.0, 178.0, f32(91.85)
    kr, kt, ks = f32(1.487), f32(0.639), f32(0.465)
    k = lat.k
    prel = 0.5 * (k[p] - k[q])
    pprel = 0.5 * (k[r] - k[s])
    qv = prel - pprel
    q2 = np.sum(qv * qv, axis=-1)
    vr = v0r * PI ** 1.5 * np.exp(-q2 / (4.0 * kr)) / (kr ** 1.5)
    vt = -v0t * PI ** 1.5 * np.exp(-q2 / (4.0 * kt)) / (kt ** 1.5)
    vs = -v0s * PI ** 1.5 * np.exp(-q2 / (4.0 * ks)) / (ks ** 1.5)
    m1, m2, m3, m4 = lat.sz[p], lat.sz[q], lat.sz[r], lat.sz[s]
    t1, t2, t3, t4 = lat.tz[p], lat.tz[q], lat.tz[r], lat.tz[s]
    ds = ((m1 == m3) & (m2 == m4)).astype(np.float64)
    dt = ((t1 == t3) & (t2 == t4)).astype(np.float64)
    ss = _sigma_dot_sigma(m1, m2, m3, m4)
    tt = _sigma_dot_sigma(t1, t2, t3, t4)
    vr = vr * (3.0 * ds * dt - tt * ds - ss * dt - tt * ss) / 8.0
    vs = vs * (3.0 * ds * dt + tt * ds - 3.0 * ss * dt - tt * ss) / 16.0
    vt = vt * (3.0 * ds * dt - 3.0 * tt * ds + ss * dt - tt * ss) / 16.0
    v = (vs + vr + vt) / lat.volume
    ok = np.all(lat.n[p] + lat.n[q] == lat.n[r] + lat.n[s], axis=-1) & (t1 + t2 == t3 + t4)
    return np.where(ok, v, 0.0)


class MinnesotaInteraction:
    def __init__(self, lat):
        self.lat = lat

    def __call__(self, p, q, r, s):
        p, q, r, s = (np.asarray(x) for x in (p, q, r, s))
        v = minnesota_direct(self.lat, p, q, r, s) - minnesota_direct(self.lat, p, q, s, r)
        return v.astype(np.complex128)


class TableInteraction:
    """Antisymmetrized elements from canonical records (p<q, r<s) of the
    reference's blocks (hhhh/hhhp/hhpp/pppp) plus full hphp records; the other
    orderings follow from antisymmetry and hermiticity."""

    def __init__(self, nsp, records):
        self.nsp = nsp
        keys, vals = [], []
        for (p, q, r, s, v) in records:
            keys.append(self._key(p, q, r, s)); vals.append(v)
        keys = np.concatenate(keys) if keys else np.zeros(0, np.int64)
        vals = np.concatenate(vals) if vals else np.zeros(0, np.complex128)
        order = np.argsort(keys, kind="stable")
        self.keys = keys[order]
        self.vals = vals[order]

    def _key(self, p, q, r, s):
        n = self.nsp
        return ((np.asarray(p, np.int64) * n + q) * n + r) * n + s

    def _lookup(self, p, q, r, s):
        key = self._key(p, q, r, s)
        idx = np.searchsorted(self.keys, key)
        idx = np.minimum(idx, len(self.keys) - 1)
        hit = self.keys[idx] == key
        return np.where(hit, self.vals[idx], 0.0), hit

    def __call__(self, p, q, r, s):
        p, q, r, s = (np.asarray(x, np.int64) for x in (p, q, r, s))
        out = np.zeros(np.broadcast(p, q, r, s).shape, np.complex128)
        found = np.zeros(out.shape, bool)
        for (a, b, c, d, sign, conj) in ((p, q, r, s, 1, False), (q, p, r, s, -1, False),
                                         (p, q, s, r, -1, False), (q, p, s, r, 1, False),
                                         (r, s, p, q, 1, True), (s, r, p, q, -1, True),
                                         (r, s, q, p, -1, True), (s, r, q, p, 1, True)):
            v, hit = self._lookup(a, b, c, d)
            if conj:
                v = np.conj(v)
            take = hit & ~found
            out = np.where(take, sign * v, out)
            found |= hit
        return out


# --------------------------------------------------------------------------- #
# channel structure (ordered pairs, reference order)
# --------------------------------------------------------------------------- #

class Channels:
    """(K, Tz) two-body channels of ordered pairs (ccm_iter.f90:223-603) and the
    particle-hole (Pandya) channels K = n_a - n_i (ccm_iter.f90:1866-2171)."""

    def __init__(self, lat: Lattice):
        self.lat = lat
        A, nsp = lat.A, lat.nsp
        holes = np.arange(A)
        parts = np.arange(A, nsp)
        self.holes, self.parts = holes, parts

        def key2(a, b):
            K = lat.n[a] + lat.n[b]
            tz = (lat.tz[a] + lat.tz[b]) // 2
            return [tuple(x) for x in np.column_stack([tz, K])]

        hh = [(i, j) for i in holes for j in holes]
        pp = [(a, b) for a in parts for b in parts]
        hh_keys = key2(np.array([x[0] for x in hh]), np.array([x[1] for x in hh]))
        pp_keys = key2(np.array([x[0] for x in pp]), np.array([x[1] for x in pp]))
        hh_by, pp_by = {}, {}
        for pr, k in zip(hh, hh_keys):
            hh_by.setdefault(k, []).append(pr)
        for pr, k in zip(pp, pp_keys):
            pp_by.setdefault(k, []).append(pr)
        keys = sorted(k for k in hh_by if k in pp_by)
        self.keys = keys
        self.hh = [np.array(hh_by[k], np.int64) for k in keys]
        self.pp = [np.array(pp_by[k], np.int64) for k in keys]
        self.hh_all = hh_by
        # pair -> (channel, row) maps; -1 when the pair is in no hhpp channel
        self.hh_ch = -np.ones((nsp, nsp), np.int64)
        self.hh_row = -np.ones((nsp, nsp), np.int64)
        self.pp_ch = -np.ones((nsp, nsp), np.int64)
        self.pp_row = -np.ones((nsp, nsp), np.int64)
        for c in range(len(keys)):
            h, p = self.hh[c], self.pp[c]
            self.hh_ch[h[:, 0], h[:, 1]] = c
            self.hh_row[h[:, 0], h[:, 1]] = np.arange(len(h))
            self.pp_ch[p[:, 0], p[:, 1]] = c
            self.pp_row[p[:, 0], p[:, 1]] = np.arange(len(p))
        self.nhh = np.array([len(x) for x in self.hh])
        self.npp = np.array([len(x) for x in self.pp])
        self.off = np.concatenate([[0], np.cumsum(self.nhh * self.npp)])
        self.size = int(self.off[-1])

        # ph channels: ph pairs (a,i) with K = n_a - n_i, hp pairs (j,b) with K = n_j - n_b
        ph = [(a, i) for a in parts for i in holes]
        hp = [(j, b) for j in holes for b in parts]

        def keyd(x, y):
            K = lat.n[x] - lat.n[y]
            tz = (lat.tz[x] - lat.tz[y]) // 2
            return [tuple(v) for v in np.column_stack([tz, K])]

        ph_keys = keyd(np.array([x[0] for x in ph]), np.array([x[1] for x in ph]))
        hp_keys = keyd(np.array([x[0] for x in hp]), np.array([x[1] for x in hp]))
        ph_by, hp_by = {}, {}
        for pr, k in zip(ph, ph_keys):
            ph_by.setdefault(k, []).append(pr)
        for pr, k in zip(hp, hp_keys):
            hp_by.setdefault(k, []).append(pr)
        self.ph_keys = sorted(k for k in ph_by if k in hp_by)
        self.ph = [np.array(ph_by[k], np.int64) for k in self.ph_keys]
        self.hp = [np.array(hp_by[k], np.int64) for k in self.ph_keys]

    def flat_index(self, a, b, i, j):
        """Flat index of ordered T(ab,ij) in the concatenated channel storage
        (row = pp pair, column = hh pair, row-major), or -1."""
        c = self.pp_ch[a, b]
        c2 = self.hh_ch[i, j]
        ok = (c >= 0) & (c == c2)
        idx = self.off[np.maximum(c, 0)] + self.pp_row[a, b] * self.nhh[np.maximum(c, 0)] + self.hh_row[i, j]
        return np.where(ok, idx, -1)


# --------------------------------------------------------------------------- #
# the CCD solver
# --------------------------------------------------------------------------- #

class CCD:
    """Ordered-pair CCD of the reference, on dense per-channel numpy blocks."""

    def __init__(self, lat: Lattice, vfun, fock=None, e0=None):
        self.lat = lat
        self.v = vfun
        self.ch = Channels(lat)
        ch = self.ch
        A = lat.A
        # V blocks (ccm_iter.f90:819-1311)
        self.v_hhpp, self.v_pppp, self.v_hhhh = [], [], []
        for c in range(len(ch.keys)):
            h, p = ch.hh[c], ch.pp[c]
            self.v_hhpp.append(vfun(h[:, 0][:, None], h[:, 1][:, None], p[:, 0][None, :], p[:, 1][None, :]))
            self.v_pppp.append(vfun(p[:, 0][:, None], p[:, 1][:, None], p[:, 0][None, :], p[:, 1][None, :]))
            self.v_hhhh.append(vfun(h[:, 0][:, None], h[:, 1][:, None], h[:, 0][None, :], h[:, 1][None, :]))
        # constant ring pieces: Vtilde((i,a),(b,j)) = <ij||ab>; W0((i,a),(j,b)) = -<ib||ja>
        self.vtil, self.w0 = [], []
        for K in range(len(ch.ph_keys)):
            ph, hp = ch.ph[K], ch.hp[K]
            i, a = hp[:, 0][:, None], hp[:, 1][:, None]
            b, j = ph[:, 0][None, :], ph[:, 1][None, :]
            self.vtil.append(vfun(i, j, a, b))
            jj, bb = hp[:, 0][None, :], hp[:, 1][None, :]
            self.w0.append(-vfun(i, bb, jj, a))
        self.fock = self._fock() if fock is None else np.asarray(fock)
        self.e0 = self._e0() if e0 is None else e0
        self.nstep_hist = []

    # ccm_main.f90:1586-1684
    def _fock(self):
        lat = self.lat
        nsp, A = lat.nsp, lat.A
        f = np.diag(lat.e.astype(np.complex128))
        h = np.arange(A)
        for p in range(nsp):
            same = np.where(np.all(lat.n == lat.n[p], 1) & (lat.tz == lat.tz[p]))[0]
            for q in same:
                f[p, q] += np.sum(self.v(np.full(A, p), h, np.full(A, q), h))
        return f

    def _e0(self):
        lat = self.lat
        A = lat.A
        i, j = np.meshgrid(np.arange(A), np.arange(A), indexing="ij")
        return np.sum(lat.e[:A]) + 0.5 * np.sum(self.v(i, j, i, j))

    # ccm_iter.f90:2329-2405
    def mbpt2_start(self):
        f = np.real(np.diag(self.fock))
        T = []
        for c in range(len(self.ch.keys)):
            h, p = self.ch.hh[c], self.ch.pp[c]
            den = (-f[p[:, 0]][:, None] - f[p[:, 1]][:, None] + f[h[:, 0]][None, :] + f[h[:, 1]][None, :])
            with np.errstate(divide="ignore", invalid="ignore"):
                t = np.where(den != 0.0, np.conj(self.v_hhpp[c]).T / den, 0.0)
            T.append(t.astype(np.complex128))
        return T

    def flat(self, T):
        return np.concatenate([t.ravel() for t in T])

    def unflat(self, x):
        ch = self.ch
        return [x[ch.off[c]:ch.off[c + 1]].reshape(ch.npp[c], ch.nhh[c]) for c in range(len(ch.keys))]

    # ccm_energy.f90:4-71
    def energy(self, T):
        return sum(0.25 * np.sum(T[c] * self.v_hhpp[c].T) for c in range(len(T)))

    def partners(self, x):
        """States with the same momentum and isospin as x (including x)."""
        cache = self.__dict__.setdefault("_partner_cache", {})
        got = cache.get(int(x))
        if got is None:
            lat = self.lat
            got = np.where(np.all(lat.n == lat.n[x], 1) & (lat.tz == lat.tz[x]))[0]
            cache[int(x)] = got
        return got

    def partner_table(self):
        """(nsp, P) table of partners(x), padded with -1."""
        pt = self.__dict__.get("_ptab")
        if pt is None:
            lists = [self.partners(x) for x in range(self.lat.nsp)]
            pt = -np.ones((self.lat.nsp, max(len(v) for v in lists)), np.int64)
            for x, v in enumerate(lists):
                pt[x, :len(v)] = v
            self._ptab = pt
        return pt

    # ccm_t2_eqn.f90:1312-1518
    def i6i7(self, T):
        lat, ch = self.lat, self.ch
        A, nsp = lat.A, lat.nsp
        pt = self.partner_table()
        I7 = np.zeros((A, A), np.complex128)
        I6 = np.zeros((nsp, nsp), np.complex128)
        for c in range(len(ch.keys)):
            h, p = ch.hh[c], ch.pp[c]
            # I7(k,i) += 0.5 * V(kl,cd) T(cd,il) for (i,l) in channel, k partner of i
            X = self.v_hhpp[c] @ T[c]          # X[(k,l),(i,l')] = sum_cd V(kl,cd) T(cd,il')
            i, l = h[:, 0], h[:, 1]
            cols = np.arange(len(h))
            for slot in range(pt.shape[1]):
                k = pt[i, slot]
                ok = (k >= 0) & (k < A)
                kk = np.where(ok, k, 0)
                ok &= ch.hh_ch[kk, l] == c
                np.add.at(I7, (k[ok], i[ok]), 0.5 * X[ch.hh_row[kk, l][ok], cols[ok]])
            # I6(a,c) += -0.5 V(kl,dc) T(da,kl) for (d,a) in channel, c partner of a
            Y = T[c] @ self.v_hhpp[c]          # Y[(d,a),(d',c)] = sum_kl T(da,kl) V(kl,d'c)
            d, a = p[:, 0], p[:, 1]
            rows = np.arange(len(p))
            for slot in range(pt.shape[1]):
                cc = pt[a, slot]
                ok = cc >= A
                ccs = np.where(ok, cc, A)
                ok &= ch.pp_ch[d, ccs] == c
                np.add.at(I6, (a[ok], cc[ok]), -0.5 * Y[rows[ok], ch.pp_row[d, ccs][ok]])
        I7 = I7 + self.fock[:A, :A]
        I6[A:, A:] = I6[A:, A:] + self.fock[A:, A:]
        return I6, I7

    # ccm_t2_eqn.f90:4-1304
    def residual(self, T, terms=15):
        """terms: bit mask of the contributions besides conj(V): 1 one-body,
        2 pp ladder, 4 hh ladder, 8 ring (for term-by-term comparisons)."""
        lat, ch = self.lat, self.ch
        A = lat.A
        nC = len(ch.keys)
        I6, I7 = self.i6i7(T)
        d_h = np.diag(I7).copy()
        d_p = np.diag(I6).copy()
        np.fill_diagonal(I7, 0.0)
        np.fill_diagonal(I6, 0.0)
        R = [np.conj(self.v_hhpp[c]).T.copy() for c in range(nC)]
        # one-body terms with P(ab) and P(ij) (ccm_t2_eqn.f90:144-211)
        pt = self.partner_table()
        for c in range(nC if terms & 1 else 0):
            h, p = ch.hh[c], ch.pp[c]
            # row (a,b): s = sum_{c' partner of b} I6(b,c') T(ac',:), added to
            # row (a,b) and subtracted from row (b,a)
            a, b = p[:, 0], p[:, 1]
            S = np.zeros((len(p), len(h)), np.complex128)
            for slot in range(pt.shape[1]):
                cc = pt[b, slot]
                ok = cc >= A
                ccs = np.where(ok, cc, A)
                ok &= ch.pp_ch[a, ccs] == c
                S[ok] += I6[b[ok], cc[ok]][:, None] * T[c][ch.pp_row[a[ok], cc[ok]], :]
            R[c] += S
            R[c][ch.pp_row[b, a], :] -= S
            # column (i,j): s = -sum_{k partner of j} I7(k,j) T(:,ik), added to
            # column (i,j) and subtracted from column (j,i)
            i, j = h[:, 0], h[:, 1]
            S = np.zeros((len(p), len(h)), np.complex128)
            for slot in range(pt.shape[1]):
                k = pt[j, slot]
                ok = (k >= 0) & (k < A)
                kk = np.where(ok, k, 0)
                ok &= ch.hh_ch[i, kk] == c
                S[:, ok] -= I7[k[ok], j[ok]][None, :] * T[c][:, ch.hh_row[i[ok], k[ok]]]
            R[c] += S
            R[c][:, ch.hh_row[j, i]] -= S
        # I9 and the two ladders (ccm_t2_eqn.f90:423-647)
        for c in range(nC):
            I9 = self.v_hhhh[c] + 0.5 * (self.v_hhpp[c] @ T[c])
            if terms & 2:
                R[c] += 0.5 * (self.v_pppp[c] @ T[c])
            if terms & 4:
                R[c] += 0.5 * (T[c] @ I9)
        if not terms & 8:
            return R, d_h, d_p
        # ring (ccm_t2_eqn.f90:37-62, 657-1185)
        Tf = self.flat(T)
        Rf = self.flat(R)
        for K in range(len(ch.ph_keys)):
            ph, hp = ch.ph[K], ch.hp[K]
            a, i = ph[:, 0][:, None], ph[:, 1][:, None]
            j, b = hp[:, 0][None, :], hp[:, 1][None, :]
            idx = ch.flat_index(a, b, i, j)
            Tt = np.where(idx >= 0, Tf[np.maximum(idx, 0)], 0.0)
            I4 = 0.5 * (self.vtil[K] @ Tt) + self.w0[K]
            Rt = Tt @ I4
            for (aa, bb, ii, jj, sg) in ((a, b, i, j, 1.0), (b, a, i, j, -1.0),
                                         (a, b, j, i, -1.0), (b, a, j, i, 1.0)):
                idx2 = ch.flat_index(aa, bb, ii, jj)
                m = idx2 >= 0
                np.add.at(Rf, idx2[m], sg * Rt[m])
        return self.unflat(Rf), d_h, d_p

    def denominators(self, d_h, d_p):
        ch = self.ch
        D = []
        for c in range(len(ch.keys)):
            h, p = ch.hh[c], ch.pp[c]
            D.append(d_h[h[:, 0]][None, :] + d_h[h[:, 1]][None, :]
                     - d_p[p[:, 0]][:, None] - d_p[p[:, 1]][:, None])
        return D

    # ccm_iter.f90:1-170 + ccm_diis.f90:74-249
    def solve(self, subspace=6, diis_step=6, tol=1e-6, maxit=100, T0=None,
              callback=None):
        T = self.mbpt2_start() if T0 is None else [t.copy() for t in T0]
        e1 = self.energy(T)
        self.mbpt2 = self.e0 + e1
        energies = [e1]
        hist = []            # t2_diis columns
        nstep = 0
        count = 0
        dener = 4.0
        e2 = 0.0
        for it in range(1, maxit + 1):
            if not abs(dener) > tol:
                continue
            R, d_h, d_p = self.residual(T)
            count += 1
            nstep += 1
            D = self.denominators(d_h, d_p)
            if nstep > subspace:
                nstep = subspace
                hist = hist[1:]
            T = [0.75 * T[c] + 0.25 * R[c] / D[c] for c in range(len(T))]
            x = self.flat(T)
            if len(hist) < nstep:
                hist.append(x)
            else:
                hist[nstep - 1] = x
            if count == diis_step:
                count = 0
                err = [hist[k + 1] - hist[k] for k in range(nstep - 1)]
                B = np.zeros((nstep, nstep), np.complex128)
                for p_ in range(nstep - 1):
                    for q_ in range(nstep - 1):
                        B[p_, q_] = np.sum(err[p_] * err[q_])
                B[: nstep - 1, nstep - 1] = -1.0
                B[nstep - 1, : nstep - 1] = -1.0
                rhs = np.zeros(nstep, np.complex128)
                rhs[nstep - 1] = -1.0
                sol = np.linalg.solve(B, rhs)
                x = sum(sol[k] * hist[k] for k in range(nstep - 1))
                T = self.unflat(x)
            e2 = self.energy(T)
            dener = e2 - e1
            e1 = e2
            energies.append(e2)
            if callback is not None:
                callback(it, e2)
        self.T = T
        self.ecc = self.e0 + e2
        return energies

    @staticmethod
    def anderson_coefficients(G):
        """min |sum_j c_j f_j| with sum_j c_j = 1 for the Gram matrix G: G c = 1
        scaled by the diagonal, then c normalized; None when a scaled pivot falls
        below 1e-13 (Gaussian elimination with partial pivoting) or the result is
        not finite.  The engine's accelerator, not a reference routine."""
        m = len(G)
        d = np.diag(G).copy()
        if not np.all(np.isfinite(d)) or np.any(d <= 0.0):
            return None
        sc = 1.0 / np.sqrt(d)
        a = sc[:, None] * G * sc[None, :]
        b = sc.copy()
        for k in range(m):
            piv = k + int(np.argmax(np.abs(a[k:, k])))
            if not abs(a[piv, k]) > 1e-13:
                return None
            if piv != k:
                a[[k, piv]] = a[[piv, k]]
                b[[k, piv]] = b[[piv, k]]
            for i in range(k + 1, m):
                f = a[i, k] / a[k, k]
                a[i, k:] -= f * a[k, k:]
                b[i] -= f * b[k]
        c = np.zeros(m)
        for i in range(m - 1, -1, -1):
            c[i] = (b[i] - a[i, i + 1:] @ c[i + 1:]) / a[i, i]
        c *= sc
        tot = c.sum()
        if not np.isfinite(tot) or tot == 0.0:
            return None
        return c / tot

    def solve_anderson(self, depth=8, beta=1.0, tol=1e-8, maxit=100, T0=None):
        """The engine's Anderson (Pulay) update (ccgpu_solve_opts.accel 1), which
        replaces the reference's diis_t1t2 schedule: with the residual map
        g(T) = R(T)/D(T) of the dynamic denominators, every iteration forms
        f = g(T) - T (kept rounded to complex64 for the Gram matrix) and
        y = T + beta f, holds the last `depth` pairs, and takes
        T = sum_j c_j y_j with the coefficients of anderson_coefficients on
        G_jk = Re<f_j, f_k>, dropping the oldest pairs while that system is
        near singular.  It stops once |dE| <= tol on two consecutive
        iterations (the reference's test, ccm_iter.f90:71-88, asks for one).
        Returns the energies (start, then one per iteration)."""
        T = self.mbpt2_start() if T0 is None else [t.copy() for t in T0]
        x = self.flat(T)
        e1 = self.energy(T)
        energies = [e1]
        ys, fs, eys = [], [], []
        small = 0
        for it in range(1, maxit + 1):
            if small >= 2:
                break
            R, d_h, d_p = self.residual(self.unflat(x))
            D = self.flat(self.denominators(d_h, d_p))
            Rf = self.flat(R)
            with np.errstate(divide="ignore", invalid="ignore"):
                f = np.where(D != 0.0, Rf / np.where(D != 0.0, D, 1.0), 0.0) - x
            f32 = f.astype(np.complex64).astype(np.complex128)
            y = x + beta * f
            if len(ys) == depth:
                ys.pop(0), fs.pop(0), eys.pop(0)
            ys.append(y)
            fs.append(f32)
            eys.append(self.energy(self.unflat(y)))
            while True:
                G = np.array([[np.real(np.vdot(fj, fk)) for fk in fs] for fj in fs])
                c = self.anderson_coefficients(G)
                if c is not None:
                    break
                if len(ys) == 1:
                    c = np.ones(1)
                    break
                ys.pop(0), fs.pop(0), eys.pop(0)
            x = sum(cj * yj for cj, yj in zip(c, ys))
            e2 = sum(cj * ej for cj, ej in zip(c, eys))
            small = small + 1 if abs(e2 - e1) <= tol else 0
            e1 = e2
            energies.append(e2)
        self.T = self.unflat(x)
        self.ecc = self.e0 + e1
        return energies

    # ccm_triples.f90:174-294, 528-735 and ccm_t3full_channels.f90:4-284
    def triples(self, T, vhppp=None, mode="t", w3=None):
        """Sums over the (K3,Tz3) channels and a<b<c, i<j<k of the triples
        amplitude numerator W(T2) (the 18 terms of pert_triples2) and D =
        f_i + f_j + f_k - f_a - f_b - f_c.  The particle term uses
        conj(<cb||kd>) as the reference does (three_body_oneparticle_1 reads
        conj(vnn_hppp(kd,cb))).

        mode "t":     |W|^2 / D, the (T) energy et_4 (ccm_triples.f90:528-735)
        mode "td":    |W + W3|^2 / D, its et_5 at tnf_approx 3 (:656-722)
        mode "ccdt1": (W3 + W) conj(W3) / D, the CCDT-1 energy of T3 =
                      (W3 + W)/D (t3_intermediate, ccm_t3_eqn.f90:1-192;
                      t3_energy, ccm_t3full_channels.f90:570-667)
        mode "w3":    |W3|^2 / D, the CCDT-1 energy of the start T3 = W3/D
                      (setup_t3, ccm_t3full_channels.f90:291-567)
        with W3 = w3(a, b, c, i, j, k), the antisymmetrized three-body element
        <abc||ijk> (chiral_3nf_asym)."""
        lat, ch = self.lat, self.ch
        A, nsp = lat.A, lat.nsp
        vh = self.v if vhppp is None else vhppp
        Tf = self.flat(T)
        f = np.diag(self.fock)
        nmax = lat.nmax
        # orbit lookup by (n, sz, tz)
        lut = {}
        for x in range(nsp):
            lut[(tuple(lat.n[x]), lat.sz[x], lat.tz[x])] = x

        def trip_key(x, y, z):
            return (lat.tz[x] + lat.tz[y] + lat.tz[z], tuple(lat.n[x] + lat.n[y] + lat.n[z]))

        hh3, pp3 = {}, {}
        for i in range(A):
            for j in range(i + 1, A):
                for k in range(j + 1, A):
                    hh3.setdefault(trip_key(i, j, k), []).append((i, j, k))
        for a in range(A, nsp):
            for b in range(a + 1, nsp):
                for c in range(b + 1, nsp):
                    key = trip_key(a, b, c)
                    if key in hh3:
                        pp3.setdefault(key, []).append((a, b, c))

        def tval(a, b, i, j):
            idx = ch.flat_index(a, b, i, j)
            return np.where(idx >= 0, Tf[np.maximum(idx, 0)], 0.0)

        def partner_list(nvec, tzv, want_particle):
            out = []
            if np.any(np.abs(nvec) > nmax) or abs(tzv) > 1:
                return out
            for szv in (-1, 1):
                x = lut.get((tuple(nvec), szv, tzv))
                if x is None:
                    continue
                if want_particle and x >= A:
                    out.append(x)
                if not want_particle and x < A:
                    out.append(x)
            return out

        def one_particle(a, b, c, i, j, k):
            s = 0.0 + 0.0j
            for d in partner_list(lat.n[i] + lat.n[j] - lat.n[a], lat.tz[i] + lat.tz[j] - lat.tz[a], True):
                if ch.pp_ch[a, d] < 0 or ch.pp_ch[a, d] != ch.hh_ch[i, j]:
                    continue
                s += np.conj(vh(c, b, k, d)) * tval(a, d, i, j)
            return s

        def one_hole(a, b, c, i, j, k):
            s = 0.0 + 0.0j
            for l in partner_list(lat.n[k] + lat.n[j] - lat.n[c], lat.tz[k] + lat.tz[j] - lat.tz[c], False):
                if ch.hh_ch[i, l] < 0 or ch.hh_ch[i, l] != ch.pp_ch[a, b]:
                    continue
                s += np.conj(self.v(j, k, l, c)) * tval(a, b, i, l)
            return s

        et = 0.0 + 0.0j
        for key, plist in pp3.items():
            for (i, j, k) in hh3[key]:
                for (a, b, c) in plist:
                    spen = f[i] + f[j] + f[k] - f[a] - f[b] - f[c]
                    x3 = 0.0 if mode == "t" else complex(w3(a, b, c, i, j, k))
                    if mode == "w3":
                        et += x3 * np.conj(x3) / spen
                        continue
                    s = (one_particle(a, b, c, i, j, k) - one_particle(b, a, c, i, j, k)
                         - one_particle(c, b, a, i, j, k) - one_particle(a, b, c, k, j, i)
                         - one_particle(a, b, c, i, k, j) + one_particle(c, b, a, i, k, j)
                         + one_particle(c, b, a, k, j, i) + one_particle(b, a, c, i, k, j)
                         + one_particle(b, a, c, k, j, i))
                    s -= (one_hole(a, b, c, i, j, k) - one_hole(a, c, b, i, j, k)
                          - one_hole(c, b, a, i, j, k) - one_hole(a, b, c, k, j, i)
                          - one_hole(a, b, c, j, i, k) + one_hole(c, b, a, j, i, k)
                          + one_hole(c, b, a, k, j, i) + one_hole(a, c, b, j, i, k)
                          + one_hole(a, c, b, k, j, i))
                    if mode == "t":
                        et += np.conj(s) * s / spen
                    elif mode == "td":
                        et += np.conj(s + x3) * (s + x3) / spen
                    else:
                        et += (x3 + s) * np.conj(x3) / spen
        return et


def load_reference_dump(path):
    """Read the binary dumps written by oracle/ref_driver.f90 (orbitals, Fock,
    E0, canonical V records, converged T2) into oracle objects."""
    import os
    with open(os.path.join(path, "orbitals.bin"), "rb") as fh:
        nsp, A = np.frombuffer(fh.read(8), np.int32)
        ints = np.frombuffer(fh.read(4 * 5 * nsp), np.int32).reshape(5, nsp)
        dbl = np.frombuffer(fh.read(8 * 4 * nsp), np.float64).reshape(4, nsp)
        box, volume, kf, rho = np.frombuffer(fh.read(32), np.float64)
    lat = Lattice(ints[0], ints[1], ints[2], ints[3], ints[4], dbl[0], dbl[1], dbl[2], dbl[3],
                  A, box, volume, kf=kf, rho=rho)
    raw = np.fromfile(os.path.join(path, "fock.bin"), np.complex128)
    fock = raw[: nsp * nsp].reshape(nsp, nsp).T      # Fortran column-major
    e0 = raw[nsp * nsp]
    rec = np.dtype([("kind", "<i4"), ("p", "<i4"), ("q", "<i4"), ("r", "<i4"), ("s", "<i4"),
                    ("v", "<c16")])
    vb = np.fromfile(os.path.join(path, "vblocks.bin"), rec)
    out = {"lat": lat, "fock": fock, "e0": e0, "vrec": vb}
    t2p = os.path.join(path, "t2.bin")
    if os.path.exists(t2p):
        trec = np.dtype([("a", "<i4"), ("b", "<i4"), ("i", "<i4"), ("j", "<i4"), ("v", "<c16")])
        out["t2"] = np.fromfile(t2p, trec)
    return out
