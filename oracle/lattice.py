"""Oracle: lattice, index maps, C4v symmetries, form factors, Fourier tables.

Restates (reference paths relative to the TUPS repository):

* IBZ bosonic index map          SRC/parquet_util.f90:110-117, 512-520
* symmetry_operation (forward)   SRC/parquet_util.f90:896-978
* my_symmetry_operation (bonds)  SRC/parquet_formfactors.f90:367-446
* inverse ops used by FF_inv     SRC/parquet_formfactors.f90:651-789
* list_symmetries                SRC/parquet_util.f90:1140-1207
* create_list / sort             SRC/parquet_formfactors.f90:594-646
* Nl from `neighbours`           SRC/parquet_formfactors.f90:494-535
* initialize_bonds               SRC/parquet_formfactors.f90:95-360
* FF_array (norm, Re+Im)         SRC/parquet_formfactors.f90:11-90
* init_FF_inv                    SRC/parquet_formfactors.f90:545-569
* init_char_table                SRC/parquet_formfactors.f90:572-588
* sym_op_L                       SRC/parquet_util.f90:732-890
* calculate_expSqR / expqSR      SRC/parquet_PhiR.f90:360-435
* NR from Nl                     SRC/preprocessing.f90:127-133
* derived sizes                  SRC/parquet_ini.f90:326-348

Pinned by (tests/test_oracle_lattice.py): exhaustive enumeration of the IBZ
map against its closed form; the star-count identity
sum_{Q in IBZ} |list(Q)| = Nx*Ny and exact coverage of the BZ by the stars;
group closure / inverse pairs of the eight C4v operations; the reference's
own closed-form check FF_sdwave_check (SRC/parquet_check.f90:264-292);
orthonormality sum_k FF_l FF_l' = delta_ll'; the transformation law
FF(S^-1 k, l) = fac(S,l) * FF(k, l_sym(S,l)) that ties FF, FF_inv, reps and
sym_op_L together; and the exponential tables against direct evaluation of
e^{i q.r} over the full BZ (sum over stars == sum over the BZ).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NS = 8  # point-group operations of the square lattice, SRC/parquet_ini.f90:26

# char_table(Ns, 5), SRC/parquet_formfactors.f90:580-586 (column-major reshape)
CHAR_TABLE = np.array(
    [
        [1, 1, 1, 1, 2],
        [1, -1, 1, -1, 0],
        [1, -1, 1, -1, 0],
        [1, 1, 1, 1, -2],
        [1, -1, -1, 1, 0],
        [1, 1, -1, -1, 0],
        [1, -1, -1, 1, 0],
        [1, 1, -1, -1, 0],
    ],
    dtype=np.int64,
)

# inverse of each operation (6 <-> 8), SRC/parquet_util.f90:906-908
SYM_INV = {1: 1, 2: 2, 3: 3, 4: 4, 5: 5, 6: 8, 7: 7, 8: 6}


def sym_vec(s: int, x: int, y: int) -> tuple[int, int]:
    """Integer C4v operation on a vector, SRC/parquet_formfactors.f90:390-444."""
    return {
        1: (x, y),
        2: (x, -y),
        3: (-x, y),
        4: (-x, -y),
        5: (y, x),
        6: (y, -x),
        7: (-y, -x),
        8: (-y, x),
    }[s]


def sym_index(s: int, ix: int, iy: int, nx: int, ny: int) -> tuple[int, int]:
    """symmetry_operation on 1-based momentum indices, SRC/parquet_util.f90:896-978."""
    x, y = sym_vec(s, ix - 1, iy - 1)
    return x % nx + 1, y % ny + 1


def list_symmetries(kx: int, ky: int, nx_ibz: int) -> list[bool]:
    """Operations that generate the star of an IBZ point, SRC/parquet_util.f90:1140-1207."""
    out = [False] * NS
    if (kx == 1 and ky == 1) or (kx == nx_ibz and ky == nx_ibz):
        out[0] = True
    elif kx == nx_ibz and ky == 1:
        out[0] = out[4] = True
    elif kx == nx_ibz and 1 < ky < nx_ibz:
        out[0] = out[1] = out[4] = out[7] = True
    elif kx > 1 and ky < nx_ibz:
        if ky == 1:
            out[0] = out[2] = out[4] = out[5] = True
        elif ky == kx:
            out[0] = out[1] = out[2] = out[3] = True
        else:
            out = [True] * NS
    return out


def create_list(nx: int) -> list[tuple[int, int, int]]:
    """Lattice sites of the IBZ sorted (stably) by |r|^2, SRC/parquet_formfactors.f90:594-646."""
    sites = [(i, j, i * i + j * j) for i in range(nx // 2 + 1) for j in range(i + 1)]
    return sorted(sites, key=lambda t: t[2])


def count_nl(nx: int, neighbours: int) -> int:
    """Number of form factors for `neighbours` shells, SRC/parquet_formfactors.f90:496-516."""
    ny = nx
    nl = 0
    for (x, y, _) in create_list(nx)[:neighbours]:
        if x == 0 and y == 0:
            nl += 1
        elif abs(x - nx / 2.0) < 0.01 and abs(y - ny / 2.0) < 0.01:
            nl += 1
        elif abs(x - nx / 2.0) < 0.01 and y == 0:
            nl += 2
        elif abs(x - nx / 2.0) < 0.01:
            nl += 4
        elif x == y or y == 0:
            nl += 4
        else:
            nl += 8
    return nl


def initialize_bonds(nx: int, nl: int):
    """First bond, character column, irrep and parity sign of each form factor.

    Follows SRC/parquet_formfactors.f90:95-360 with the default pick list
    (first Nl candidates, :522-528).  Returns a list of tuples
    (bx, by, char_col, rep, sign).
    """
    ny = nx
    bonds = []
    for (x, y, _) in create_list(nx):
        if x == 0 and y == 0:
            cands = [(x, y, 1, 1, 1)]
        elif abs(x - nx / 2.0) < 0.01 and abs(y - ny / 2.0) < 0.01:
            cands = [(x, y, 1, 1, 1)]
        elif abs(x - nx / 2.0) < 0.01 and y == 0:
            cands = [(x, y, 1, 1, 1), (x, y, 3, 3, 1)]
        elif x == y:
            cands = [(x, y, 1, 1, 1), (x, y, 4, 4, 1), (x, y, 5, 5, -1), (-y, x, 5, 6, -1)]
        elif y == 0 or abs(x - nx / 2.0) < 0.01:
            cands = [(x, y, 1, 1, 1), (x, y, 3, 3, 1), (x, y, 5, 5, -1), (-y, x, 5, 6, -1)]
        else:
            cands = [
                (x, y, 1, 1, 1), (x, y, 2, 2, 1), (x, y, 3, 3, 1), (x, y, 4, 4, 1),
                (x, y, 5, 5, -1), (x, -y, 5, 6, -1), (y, x, 5, 5, -1), (-y, x, 5, 6, -1),
            ]
        for c in cands:
            bonds.append(c)
            if len(bonds) == nl:
                return bonds
    raise ValueError(f"lattice Nx={nx} has fewer than Nl={nl} form factors")


def form_factors(nx: int, nl: int):
    """FF(kx, ky, l), signs(l), reps(l); SRC/parquet_formfactors.f90:11-90.

    Returns FF as float64 array [nx, ny, nl] indexed with 0-based (kx-1, ky-1, l-1).
    """
    ny = nx
    bonds = initialize_bonds(nx, nl)
    kx = np.arange(nx)
    ky = np.arange(ny)
    px = (2.0 / nx * math.pi * kx)[:, None]
    py = (2.0 / ny * math.pi * ky)[None, :]
    ff = np.zeros((nx, ny, nl))
    for l, (bx, by, col, _rep, _sign) in enumerate(bonds):
        comp = np.zeros((nx, ny), dtype=np.complex128)
        for s in range(1, NS + 1):
            sx, sy = sym_vec(s, bx, by)
            comp += CHAR_TABLE[s - 1, col - 1] * np.exp(1j * (sx * px + sy * py))
        norm = np.sum(np.abs(comp) * np.abs(comp))
        ff[:, :, l] = 1.0 / math.sqrt(norm) * comp.real + 1.0 / math.sqrt(norm) * comp.imag
    signs = np.array([b[4] for b in bonds], dtype=np.int64)
    reps = np.array([b[3] for b in bonds], dtype=np.int64)
    return ff, signs, reps


def ff_inverse(ff: np.ndarray) -> np.ndarray:
    """FF_inv(s, kx, ky, l) = FF(S_s^-1 k, l); SRC/parquet_formfactors.f90:545-569, 651-789.

    Returns float64 array [NS, nx, ny, nl] (0-based).
    """
    nx, ny, nl = ff.shape
    out = np.zeros((NS, nx, ny, nl))
    for s in range(1, NS + 1):
        for ix in range(1, nx + 1):
            for iy in range(1, ny + 1):
                jx, jy = sym_index(SYM_INV[s], ix, iy, nx, ny)
                out[s - 1, ix - 1, iy - 1, :] = ff[jx - 1, jy - 1, :]
    return out


_SYM_L5_LOW = {1: 1, 2: 1, 3: -1, 4: -1, 5: 1, 6: 1, 7: -1, 8: -1}
_SYM_L5_HIGH = {1: (1, 0), 2: (-1, 1), 3: (1, 1), 4: (-1, 0), 5: (1, 0), 6: (1, 1), 7: (-1, 0), 8: (-1, 1)}
_SYM_L6_LOW = {1: 1, 2: -1, 3: 1, 4: -1, 5: 1, 6: -1, 7: -1, 8: 1}
_SYM_L6_HIGH = {1: (1, 0), 2: (-1, -1), 3: (1, -1), 4: (-1, 0), 5: (-1, 0), 6: (-1, -1), 7: (1, 0), 8: (1, -1)}


def sym_op_L(s: int, l: int, reps: np.ndarray) -> tuple[int, int]:
    """(l_sym, fac) of an l-index under operation s (1-based), SRC/parquet_util.f90:732-890."""
    rep = int(reps[l - 1])
    if 1 <= rep <= 4:
        return l, int(CHAR_TABLE[s - 1, rep - 1])
    if rep == 5:
        if l <= 5:
            return (l if s <= 4 else l + 1), _SYM_L5_LOW[s]
        fac, dl = _SYM_L5_HIGH[s]
        return l + dl, fac
    if rep == 6:
        if l <= 5:
            return (l if s <= 4 else l - 1), _SYM_L6_LOW[s]
        fac, dl = _SYM_L6_HIGH[s]
        return l + dl, fac
    raise ValueError(f"unknown representation {rep}")


def sym_l_tables(reps: np.ndarray):
    """Arrays lsym[s, l] (0-based l) and lfac[s, l] for all operations."""
    nl = len(reps)
    lsym = np.zeros((NS, nl), dtype=np.int64)
    lfac = np.zeros((NS, nl), dtype=np.int64)
    for s in range(1, NS + 1):
        for l in range(1, nl + 1):
            ls, f = sym_op_L(s, l, reps)
            lsym[s - 1, l - 1] = ls - 1
            lfac[s - 1, l - 1] = f
    return lsym, lfac


def nr_for_nl(nl: int) -> int:
    """NR chosen from Nl, SRC/preprocessing.f90:127-133."""
    return {1: 1, 5: 4, 9: 6}[nl]


@dataclass
class Lattice:
    """Sizes and tables of one TUPS configuration (SRC/parquet_ini.f90:326-348)."""

    nx: int
    nf: int
    nl: int
    nr: int | None = None
    ny: int = field(init=False)
    nx_ibz: int = field(init=False)
    nibz: int = field(init=False)
    nw: int = field(init=False)
    nq: int = field(init=False)
    nc: int = field(init=False)
    nz: int = field(init=False)

    def __post_init__(self):
        if self.nx % 2 or self.nf % 2:
            raise ValueError("Nx and Nf must be even (list_symmetries, bosonic box)")
        self.ny = self.nx
        self.nx_ibz = (self.nx + 2) // 2
        self.nibz = self.nx_ibz * (self.nx_ibz + 1) // 2
        self.nw = self.nf // 2
        self.nq = self.nibz * self.nw
        self.nc = self.nx * self.ny
        self.nz = self.nl * self.nf
        if self.nr is None:
            self.nr = nr_for_nl(self.nl)
        if self.nr > self.nibz:
            raise ValueError("NR exceeds the number of IBZ points")
        self.ff, self.signs, self.reps = form_factors(self.nx, self.nl)
        self.ff_inv = ff_inverse(self.ff)
        self.lsym, self.lfac = sym_l_tables(self.reps)
        self.ibz_points = self._ibz_points()
        self.qx, self.qy, self.qw = self._q_map()
        self.sym_list = np.array(
            [list_symmetries(kx, ky, self.nx_ibz) for (kx, ky) in self.ibz_points], dtype=bool
        )
        self.expSqR = calculate_expSqR(self)
        self.expqSR = calculate_expqSR(self)

    def _ibz_points(self):
        pts = []
        for kx in range(1, self.nx_ibz + 1):
            for ky in range(1, kx + 1):
                pts.append((kx, ky))
        return pts

    def _q_map(self):
        """Index_bosonic_IBZ, SRC/parquet_util.f90:110-117 (1-based ix, iy, iw)."""
        qx = np.zeros(self.nq, dtype=np.int64)
        qy = np.zeros(self.nq, dtype=np.int64)
        qw = np.zeros(self.nq, dtype=np.int64)
        for kx in range(1, self.nx_ibz + 1):
            for ky in range(1, kx + 1):
                for w in range(1, self.nw + 1):
                    idx = (((kx - 1) * kx) // 2 + ky - 1) * self.nw + w
                    qx[idx - 1], qy[idx - 1], qw[idx - 1] = kx, ky, w
        return qx, qy, qw

    def q_momentum_index(self, q0: int) -> int:
        """0-based IBZ momentum slot of 0-based q index."""
        return q0 // self.nw


def calculate_expSqR(lat: Lattice) -> np.ndarray:
    """expSqR(Q, R, sym) = e^{+i (S Q).R} on the star, else 0; SRC/parquet_PhiR.f90:360-395."""
    out = np.zeros((lat.nibz, lat.nr, NS), dtype=np.complex128)
    for m, (qx, qy) in enumerate(lat.ibz_points):
        sl = list_symmetries(qx, qy, lat.nx_ibz)
        for s in range(1, NS + 1):
            if not sl[s - 1]:
                continue
            sx, sy = sym_index(s, qx, qy, lat.nx, lat.ny)
            for r in range(lat.nr):
                rx, ry = lat.ibz_points[r]
                qdotr = (2.0 * math.pi / lat.nx) * (sx - 1) * (rx - 1) + (2.0 * math.pi / lat.ny) * (sy - 1) * (ry - 1)
                out[m, r, s - 1] = complex(math.cos(qdotr), math.sin(qdotr))
    return out


def calculate_expqSR(lat: Lattice) -> np.ndarray:
    """expqSR(k, R, sym) = e^{-i k.(S R)} on the star of R, else 0; SRC/parquet_PhiR.f90:400-435.

    k index = (ky-1)*Nx + kx (x fastest), returned 0-based as [Nc, NR, NS].
    """
    out = np.zeros((lat.nc, lat.nr, NS), dtype=np.complex128)
    for r in range(lat.nr):
        rx, ry = lat.ibz_points[r]
        sl = list_symmetries(rx, ry, lat.nx_ibz)
        for s in range(1, NS + 1):
            if not sl[s - 1]:
                continue
            srx, sry = sym_index(s, rx, ry, lat.nx, lat.ny)
            for ky in range(1, lat.ny + 1):
                for kx in range(1, lat.nx + 1):
                    k = (ky - 1) * lat.nx + kx - 1
                    kdot = (2.0 * math.pi / lat.nx) * (kx - 1) * (srx - 1) + (2.0 * math.pi / lat.ny) * (ky - 1) * (sry - 1)
                    out[k, r, s - 1] = complex(math.cos(kdot), -math.sin(kdot))
    return out


# --- frequency helpers, SRC/parquet_EQContribs.f90:17-39, 531-585 -------------

def mF(nu, nf):
    return -nu + nf + 1


def FpB(nu, om):
    return nu + om - 1


def BpBmF(o1, o2, nu, nf):
    return o1 + o2 - 1 + (-nu + nf + 1) - 1


def FmB(nu, om):
    return nu - om + 1


def BmBmF(o1, o2, nu, nf):
    return (o1 - 1) - (nu + o2 - 1) + nf + 1


def pick_nu(x, nf):
    """pickNu(x + 2*Nf): clamp of x to [1, Nf] (array of length 5*Nf), :17-39."""
    idx = x + 2 * nf
    if idx <= 2 * nf:
        return 1
    if 2 * nf < idx < 3 * nf:
        return idx - 2 * nf
    return nf


def matsubara(nu, nf, beta):
    """get_freq: i*pi/beta*(2*(nu-Nf/2-1)+1), SRC/parquet_util.f90:1332-1339."""
    return 1j * (math.pi / beta * (2.0 * (nu - nf // 2 - 1) + 1.0))
